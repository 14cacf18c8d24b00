// The hip platform's classes for the reference's own C++ tree (see HipFactory.h).
// Each method forwards to the C ABI of libfts_hip.so; the propagator plan handed to
// fts_solver_create is the reference's PropagatorAnalyzer output and Scheduler
// schedule (PropagatorAnalyzer.cpp:142-269, Scheduler.cpp:11-190) in the flat form
// of fts_plan, with propagators in ComparePropagatorKey order, blocks in the order of
// the analyzer's block map and monomers in name order.
#include "HipFactory.h"

#include <algorithm>
#include <cmath>
#include <iostream>
#include <memory>
#include <tuple>

#include "PropagatorCode.h"
#include "Scheduler.h"

namespace {

void check(int rc) {
    if (rc != FTS_OK) throw_with_line_number(std::string(fts_last_error()));
}

int* ip(std::vector<int>& v) {
    if (v.empty()) v.push_back(0);
    return v.data();
}

double* dp(std::vector<double>& v) {
    if (v.empty()) v.push_back(0.0);
    return v.data();
}

}  // namespace

fts_context* hip_platform_context() {
    static std::unique_ptr<fts_context, int (*)(fts_context*)> ctx(nullptr, fts_context_destroy);
    if (!ctx) {
        fts_context* c = nullptr;
        check(fts_context_create(0, &c));
        ctx.reset(c);
    }
    return ctx.get();
}

// ---------------------------------------------------------------- Array (host arithmetic)
HipArray::HipArray(unsigned int size) : data_(size, 0.0) { this->size = size; this->device = "cpu"; }
HipArray::HipArray(double* data, unsigned int size) : data_(data, data + size) {
    this->size = size;
    this->device = "cpu";
}
void HipArray::add(const Array& a, const Array& b) {
    for (unsigned i = 0; i < size; i++) data_[i] = a[i] + b[i];
}
void HipArray::subtract(const Array& a, const Array& b) {
    for (unsigned i = 0; i < size; i++) data_[i] = a[i] - b[i];
}
void HipArray::multiply(const Array& a, const Array& b) {
    for (unsigned i = 0; i < size; i++) data_[i] = a[i] * b[i];
}
void HipArray::divide(const Array& a, const Array& b) {
    for (unsigned i = 0; i < size; i++) data_[i] = a[i] / b[i];
}
void HipArray::linear_scaling(const Array& src, const double a, const double b) {
    for (unsigned i = 0; i < size; i++) data_[i] = a * src[i] + b;
}
void HipArray::operator=(const Array& src) {
    if (src.get_size() != size) throw_with_line_number("Sizes of arrays do not match.");
    for (unsigned i = 0; i < size; i++) data_[i] = src[i];
}
void HipArray::set_data(double* data, unsigned int n) {
    if (n != size) throw_with_line_number("Sizes of arrays do not match.");
    std::copy(data, data + n, data_.begin());
}
std::vector<double> HipArray::to_vector() const { return data_; }
double* HipArray::get_ptr() const { return const_cast<double*>(data_.data()); }
unsigned int HipArray::get_size() const { return size; }
double HipArray::operator[](unsigned int i) const { return data_[i]; }

// ---------------------------------------------------------------- ComputationBox
HipComputationBox::HipComputationBox(std::vector<int> nx, std::vector<double> lx, std::vector<std::string> bc,
                                     const double* mask)
    : ComputationBox(nx, lx, bc, mask) {
    check(fts_grid_create(hip_platform_context(), get_dim(), nx.data(), lx.data(), mask, &g_));
}
HipComputationBox::~HipComputationBox() {
    if (g_) fts_grid_destroy(g_);
}
void HipComputationBox::set_lx(std::vector<double> new_lx) {
    ComputationBox::set_lx(new_lx);
    check(fts_grid_set_lx(g_, new_lx.data()));
}
double HipComputationBox::integral_device(const double* g) {
    double v;
    check(fts_grid_integral(g_, g, 1, &v));
    return v;
}
double HipComputationBox::inner_product_device(const double* g, const double* h) {
    double v;
    check(fts_grid_inner_product(g_, g, h, 1, &v));
    return v;
}
double HipComputationBox::inner_product_inverse_weight_device(const double*, const double*, const double*) {
    throw_with_line_number("inner_product_inverse_weight_device: use the host form (ComputationBox.cpp:280-288)");
}
double HipComputationBox::multi_inner_product_device(int n_comp, const double* g, const double* h) {
    double s = 0.0;
    for (int c = 0; c < n_comp; c++) s += inner_product_device(g + (size_t)c * n_grid, h + (size_t)c * n_grid);
    return s;
}
void HipComputationBox::zero_mean_device(double*) {
    throw_with_line_number("zero_mean_device: use the host form (ComputationBox.cpp:300-308)");
}

// ---------------------------------------------------------------- PropagatorComputation
HipPropagatorComputation::HipPropagatorComputation(HipComputationBox* cb, Molecules* molecules,
                                                   PropagatorAnalyzer* analyzer, bool realspace)
    : PropagatorComputation(cb, molecules, analyzer), hcb_(cb) {
    auto& props = analyzer->get_computation_propagators();
    if (props.empty()) throw_with_line_number("There is no propagator code. Add polymers first.");
    auto& blocks = analyzer->get_computation_blocks();
    std::map<std::string, int> mon;
    for (const auto& b : molecules->get_bond_lengths()) {
        mon[b.first] = (int)monomers_.size();
        monomers_.push_back(b.first);
    }
    std::vector<double> seg;
    for (const auto& m : monomers_) seg.push_back(molecules->get_bond_lengths().at(m));
    std::vector<int> p_mon, p_max, p_kind, p_qi, p_doff, p_nd, d_prop, d_n, d_rep, p_joff, p_nj, je;
    for (const auto& it : props) prop_index_[it.first] = (int)prop_index_.size();
    for (const auto& it : props) {
        const std::string& key = it.first;
        const ComputationEdge& e = it.second;
        p_mon.push_back(mon.at(e.monomer_type));
        p_max.push_back(e.max_n_segment);
        p_doff.push_back((int)d_prop.size());
        p_nd.push_back((int)e.deps.size());
        for (const auto& d : e.deps) {
            d_prop.push_back(prop_index_.at(std::get<0>(d)));
            d_n.push_back(std::get<1>(d));
            d_rep.push_back(std::get<2>(d));
        }
        p_joff.push_back((int)je.size());
        p_nj.push_back((int)e.junction_ends.size());
        for (int n : e.junction_ends) je.push_back(n);
        if (e.deps.empty()) {
            if (key[0] == '{') {
                const std::string name = PropagatorCode::get_q_input_idx_from_key(key);
                auto f = std::find(qinit_names_.begin(), qinit_names_.end(), name);
                if (f == qinit_names_.end()) f = qinit_names_.insert(qinit_names_.end(), name);
                p_kind.push_back(FTS_LEAF_QINIT);
                p_qi.push_back((int)(f - qinit_names_.begin()));
            } else {
                p_kind.push_back(FTS_LEAF);
                p_qi.push_back(-1);
            }
        } else {
            p_kind.push_back(key[0] == '[' ? FTS_AGGREGATED : FTS_JUNCTION);
            p_qi.push_back(-1);
        }
    }
    std::vector<int> b_pol, b_l, b_r, b_nl, b_nr, b_rep, b_mon, b_nvu;
    for (const auto& it : blocks) {
        block_keys_.push_back(it.first);
        b_pol.push_back(std::get<0>(it.first));
        b_l.push_back(prop_index_.at(std::get<1>(it.first)));
        b_r.push_back(prop_index_.at(std::get<2>(it.first)));
        b_nl.push_back(it.second.n_segment_left);
        b_nr.push_back(it.second.n_segment_right);
        b_rep.push_back(it.second.n_repeated);
        b_mon.push_back(mon.at(it.second.monomer_type));
        b_nvu.push_back((int)it.second.v_u.size());
    }
    std::vector<double> vf, alpha, svf;
    std::vector<int> smon;
    for (int p = 0; p < molecules->get_n_polymer_types(); p++) {
        vf.push_back(molecules->get_polymer(p).get_volume_fraction());
        alpha.push_back(molecules->get_polymer(p).get_alpha());
    }
    for (int s = 0; s < molecules->get_n_solvent_types(); s++) {
        svf.push_back(std::get<0>(molecules->get_solvent(s)));
        smon.push_back(mon.at(std::get<1>(molecules->get_solvent(s))));
    }
    // the reference's own scheduler (8 lanes: the concurrency the device batches)
    Scheduler sc(props, 8);
    std::vector<int> s_off{0}, j_prop, j_from, j_to;
    for (const auto& slice : sc.get_schedule()) {
        for (const auto& j : slice) {
            j_prop.push_back(prop_index_.at(std::get<0>(j)));
            j_from.push_back(std::get<1>(j));
            j_to.push_back(std::get<2>(j));
        }
        s_off.push_back((int)j_prop.size());
    }
    fts_plan pl{};
    pl.model = molecules->get_model_name() == "continuous" ? FTS_CONTINUOUS : FTS_DISCRETE;
    pl.ds = molecules->get_ds();
    pl.n_monomers = (int)monomers_.size();
    pl.segment_length = dp(seg);
    pl.n_props = (int)props.size();
    pl.prop_monomer = ip(p_mon);
    pl.prop_max_n = ip(p_max);
    pl.prop_kind = ip(p_kind);
    pl.prop_qinit = ip(p_qi);
    pl.prop_dep_off = ip(p_doff);
    pl.prop_n_deps = ip(p_nd);
    pl.dep_prop = ip(d_prop);
    pl.dep_n = ip(d_n);
    pl.dep_rep = ip(d_rep);
    pl.prop_je_off = ip(p_joff);
    pl.prop_n_je = ip(p_nj);
    pl.je = ip(je);
    pl.n_blocks = (int)blocks.size();
    pl.blk_polymer = ip(b_pol);
    pl.blk_left = ip(b_l);
    pl.blk_right = ip(b_r);
    pl.blk_n_left = ip(b_nl);
    pl.blk_n_right = ip(b_nr);
    pl.blk_rep = ip(b_rep);
    pl.blk_monomer = ip(b_mon);
    pl.blk_n_vu = ip(b_nvu);
    pl.n_polymers = molecules->get_n_polymer_types();
    pl.pol_vf = dp(vf);
    pl.pol_alpha = dp(alpha);
    pl.n_solvents = molecules->get_n_solvent_types();
    pl.sol_vf = dp(svf);
    pl.sol_monomer = ip(smon);
    pl.n_qinit = (int)qinit_names_.size();
    pl.n_slices = (int)sc.get_schedule().size();
    pl.slice_off = ip(s_off);
    pl.job_prop = ip(j_prop);
    pl.job_from = ip(j_from);
    pl.job_to = ip(j_to);
    if (realspace) {
        std::vector<int> bc;
        for (const auto& b : cb->get_boundary_conditions())
            bc.push_back(b == BoundaryCondition::PERIODIC     ? FTS_BC_PERIODIC
                         : b == BoundaryCondition::REFLECTING ? FTS_BC_REFLECTING
                                                              : FTS_BC_ABSORBING);
        check(fts_solver_create_realspace(cb->handle(), &pl, bc.data(), &s_));
    } else {
        check(fts_solver_create(cb->handle(), &pl, FTS_FP64, &s_));
    }
}

HipPropagatorComputation::~HipPropagatorComputation() {
    if (s_) fts_solver_destroy(s_);
}

void HipPropagatorComputation::update_laplacian_operator() { check(fts_solver_update_laplacian(s_)); }

void HipPropagatorComputation::compute_propagators(std::map<std::string, const double*> w_block,
                                                   std::map<std::string, const double*> q_init) {
    const int M = cb->get_n_grid();
    static std::vector<double> zeros;
    std::vector<const double*> w;
    for (const auto& m : monomers_) {
        auto it = w_block.find(m);
        if (it != w_block.end()) {
            w.push_back(it->second);
        } else {
            if ((int)zeros.size() < M) zeros.assign(M, 0.0);
            w.push_back(zeros.data());
        }
    }
    last_w_.clear();
    for (const auto& it : w_block) last_w_[it.first].assign(it.second, it.second + M);
    std::vector<const double*> q;
    for (const auto& name : qinit_names_) {
        auto it = q_init.find(name);
        if (it == q_init.end()) throw_with_line_number("Could not find q_init[\"" + name + "\"].");
        q.push_back(it->second);
    }
    check(fts_solver_compute_propagators(s_, w.data(), q.empty() ? nullptr : q.data(), 0));
    partitions();
}

void HipPropagatorComputation::partitions() {
    for (int p = 0; p < molecules->get_n_polymer_types(); p++)
        check(fts_solver_get_partition(s_, p, &single_polymer_partitions[p]));
    for (int s = 0; s < molecules->get_n_solvent_types(); s++)
        check(fts_solver_get_solvent_partition(s_, s, &single_solvent_partitions[s]));
}

void HipPropagatorComputation::compute_concentrations() { check(fts_solver_compute_concentrations(s_)); }

void HipPropagatorComputation::compute_statistics(std::map<std::string, const double*> w_block,
                                                  std::map<std::string, const double*> q_init) {
    compute_propagators(w_block, q_init);
    compute_concentrations();
}

void HipPropagatorComputation::compute_stress() {
    check(fts_solver_compute_stress(s_));
    for (int p = 0; p < molecules->get_n_polymer_types(); p++) {
        double d[3] = {0.0, 0.0, 0.0};
        check(fts_solver_get_dq_dl(s_, p, d));
        dq_dl[p] = {d[0], d[1], d[2]};
    }
}

double HipPropagatorComputation::get_total_partition(int polymer) { return single_polymer_partitions[polymer]; }

void HipPropagatorComputation::get_chain_propagator(double* q_out, int polymer, int v, int u, int n) {
    const std::string key = molecules->get_polymer(polymer).get_propagator_key(v, u);
    auto it = prop_index_.find(key);
    if (it == prop_index_.end())
        throw_with_line_number("Could not find the propagator code '" + key +
                               "'. Disable 'aggregation' option to obtain propagator_analyzer.");
    check(fts_solver_get_propagator(s_, it->second, n, 0, q_out, 0));
}

// the block (and solvent) terms of one monomer type (CpuComputationContinuous.cpp:477-545)
void HipPropagatorComputation::sum_blocks(int polymer, const std::string& monomer, bool with_solvents, double scale,
                                          double* phi) {
    std::vector<int> kinds, idx;
    std::vector<double> sc;
    auto& blocks = propagator_analyzer->get_computation_blocks();
    for (size_t b = 0; b < block_keys_.size(); b++) {
        const auto& k = block_keys_[b];
        if (polymer >= 0 && std::get<0>(k) != polymer) continue;
        if (PropagatorCode::get_monomer_type_from_key(std::get<1>(k)) == monomer &&
            blocks.at(k).n_segment_right != 0) {
            kinds.push_back(FTS_SRC_BLOCK);
            idx.push_back((int)b);
            sc.push_back(scale);
        }
    }
    if (with_solvents)
        for (int s = 0; s < molecules->get_n_solvent_types(); s++)
            if (std::get<1>(molecules->get_solvent(s)) == monomer) {
                kinds.push_back(FTS_SRC_SOLVENT);
                idx.push_back(s);
                sc.push_back(1.0);
            }
    check(fts_solver_sum_concentrations(s_, (int)kinds.size(), ip(kinds), ip(idx), dp(sc), phi, 0));
}

void HipPropagatorComputation::get_total_concentration(std::string monomer_type, double* phi) {
    sum_blocks(-1, monomer_type, true, 1.0, phi);
}

void HipPropagatorComputation::get_total_concentration(int polymer, std::string monomer_type, double* phi) {
    sum_blocks(polymer, monomer_type, false, 1.0, phi);
}

void HipPropagatorComputation::get_total_concentration_gce(double fugacity, int polymer, std::string monomer_type,
                                                           double* phi) {
    Polymer& pc = molecules->get_polymer(polymer);
    const double norm = fugacity / pc.get_volume_fraction() * pc.get_alpha() * single_polymer_partitions[polymer];
    sum_blocks(polymer, monomer_type, false, norm, phi);
}

void HipPropagatorComputation::get_block_concentration(int polymer, double* phi) {
    if (propagator_analyzer->is_aggregated())
        throw_with_line_number("Disable 'aggregation' option to invoke 'get_block_concentration'.");
    Polymer& pc = molecules->get_polymer(polymer);
    const int M = cb->get_n_grid();
    auto& pblocks = pc.get_blocks();
    for (size_t b = 0; b < pblocks.size(); b++) {
        std::string kl = pc.get_propagator_key(pblocks[b].v, pblocks[b].u);
        std::string kr = pc.get_propagator_key(pblocks[b].u, pblocks[b].v);
        if (kl < kr) kl.swap(kr);
        const auto it = std::find(block_keys_.begin(), block_keys_.end(), std::make_tuple(polymer, kl, kr));
        if (it == block_keys_.end()) throw_with_line_number("block (" + kl + ", " + kr + ") not found");
        const int kind = FTS_SRC_BLOCK, i = (int)(it - block_keys_.begin());
        const double one = 1.0;
        check(fts_solver_sum_concentrations(s_, 1, &kind, &i, &one, phi + b * (size_t)M, 0));
    }
}

double HipPropagatorComputation::get_solvent_partition(int s) { return single_solvent_partitions[s]; }

void HipPropagatorComputation::get_solvent_concentration(int s, double* phi) {
    const int kind = FTS_SRC_SOLVENT;
    const double one = 1.0;
    check(fts_solver_sum_concentrations(s_, 1, &kind, &s, &one, phi, 0));
}

// CpuComputationContinuous.cpp:755-817 / CpuComputationDiscrete.cpp:1027-1089
bool HipPropagatorComputation::check_total_partition() {
    const int M = cb->get_n_grid();
    const bool continuous = molecules->get_model_name() == "continuous";
    std::vector<std::vector<double>> totals(molecules->get_n_polymer_types());
    std::vector<double> a(M), b(M), inv_e(M);
    auto& blocks = propagator_analyzer->get_computation_blocks();
    for (const auto& k : block_keys_) {
        const ComputationBlock& blk = blocks.at(k);
        const int L = prop_index_.at(std::get<1>(k)), R = prop_index_.at(std::get<2>(k));
        const int nl = blk.n_segment_left, nr = blk.n_segment_right;
        if (!continuous) {
            const std::vector<double>& w = last_w_.at(blk.monomer_type);
            for (int i = 0; i < M; i++) inv_e[i] = std::exp(w[i] * molecules->get_ds());
        }
        for (int n = continuous ? 0 : 1; n <= nr; n++) {
            check(fts_solver_get_propagator(s_, L, continuous ? nl - n : nl - n + 1, 0, a.data(), 0));
            check(fts_solver_get_propagator(s_, R, n, 0, b.data(), 0));
            if (!continuous)
                for (int i = 0; i < M; i++) b[i] *= inv_e[i];
            totals[std::get<0>(k)].push_back(cb->inner_product(a.data(), b.data()) * blk.n_repeated /
                                             cb->get_volume() / (double)blk.v_u.size());
        }
    }
    std::cout << "Polymer id: maximum,  minimum, and difference of total partitions" << std::endl;
    bool ok = true;
    for (size_t p = 0; p < totals.size(); p++) {
        const auto mm = std::minmax_element(totals[p].begin(), totals[p].end());
        std::cout << "\t" << p << ": " << *mm.second << ", " << *mm.first << ", " << *mm.second - *mm.first
                  << std::endl;
        if (*mm.second - *mm.first > 1e-7) ok = false;
    }
    return ok;
}

// ---------------------------------------------------------------- Anderson mixing
HipAndersonMixing::HipAndersonMixing(int n_var, int max_hist, double start_error, double mix_min, double mix_init)
    : AndersonMixing(n_var, max_hist, start_error, mix_min, mix_init) {
    check(fts_am_create(hip_platform_context(), n_var, max_hist, start_error, mix_min, mix_init, &am_));
}
HipAndersonMixing::~HipAndersonMixing() {
    if (am_) fts_am_destroy(am_);
}
void HipAndersonMixing::reset_count() { check(fts_am_reset(am_)); }
void HipAndersonMixing::calculate_new_fields(double* w_new, double* w_current, double* w_deriv,
                                             double old_error_level, double error_level) {
    check(fts_am_calculate_new_fields(am_, w_new, w_current, w_deriv, old_error_level, error_level, 0));
}

// ---------------------------------------------------------------- factory
HipFactory::HipFactory(std::string chain_model, bool reduce_memory_usage) {
    this->chain_model = chain_model;
    this->reduce_memory_usage = reduce_memory_usage;
}
Array* HipFactory::create_array(unsigned int size) { return new HipArray(size); }
Array* HipFactory::create_array(double* data, unsigned int size) { return new HipArray(data, size); }
ComputationBox* HipFactory::create_computation_box(std::vector<int> nx, std::vector<double> lx,
                                                   std::vector<std::string> bc, const double* mask) {
    return new HipComputationBox(nx, lx, bc, mask);
}
Molecules* HipFactory::create_molecules_information(std::string chain_model, double ds,
                                                    std::map<std::string, double> bond_lengths) {
    return new Molecules(chain_model, ds, bond_lengths);
}
PropagatorComputation* HipFactory::create_pseudospectral_solver(ComputationBox* cb, Molecules* molecules,
                                                                PropagatorAnalyzer* analyzer) {
    auto* hcb = dynamic_cast<HipComputationBox*>(cb);
    if (!hcb) throw_with_line_number("the hip platform needs a box made by its own factory");
    return new HipPropagatorComputation(hcb, molecules, analyzer, false);
}
PropagatorComputation* HipFactory::create_realspace_solver(ComputationBox* cb, Molecules* molecules,
                                                           PropagatorAnalyzer* analyzer) {
    auto* hcb = dynamic_cast<HipComputationBox*>(cb);
    if (!hcb) throw_with_line_number("the hip platform needs a box made by its own factory");
    return new HipPropagatorComputation(hcb, molecules, analyzer, true);
}
AndersonMixing* HipFactory::create_anderson_mixing(int n_var, int max_hist, double start_error, double mix_min,
                                                   double mix_init) {
    return new HipAndersonMixing(n_var, max_hist, start_error, mix_min, mix_init);
}
void HipFactory::display_info() {
    char buf[512];
    check(fts_context_info(hip_platform_context(), buf, sizeof(buf)));
    std::cout << "========== HIP platform (hip) ==========" << std::endl << buf << std::endl;
}
