// The "hip" platform: reference-shaped host objects over the C ABI of libfts_hip.so.
#include "fts_platform.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <mutex>

namespace lfts {

void check_fts(int rc) {
    if (rc != FTS_OK) throw std::runtime_error(fts_last_error());
}

fts_context* hip_context() {
    static fts_context* ctx = nullptr;
    static std::once_flag once;
    std::call_once(once, [] {
        const char* env = std::getenv("LFTS_HIP_DEVICE");
        const int dev = env ? std::atoi(env) : 0;
        check_fts(fts_context_create(dev, &ctx));
    });
    return ctx;
}

// ------------------------------------------------------------------ ComputationBox
// ComputationBox.cpp:10-157
ComputationBox::ComputationBox(std::vector<int> nx, std::vector<double> lx, std::vector<std::string> bc,
                               const double* mask) {
    if (nx.size() != lx.size())
        LFTS_THROW("The sizes of nx (" + std::to_string(nx.size()) + ") and lx (" + std::to_string(lx.size()) +
                   ") must match.");
    if (nx.size() < 1 || nx.size() > 3) LFTS_THROW("We expect 1D, 2D or 3D, but we get " + std::to_string(nx.size()));
    for (int n : nx)
        if (n <= 0) LFTS_THROW("nx must be positive numbers");
    for (double l : lx)
        if (l <= 0.0) LFTS_THROW("lx must be positive numbers");
    nx_ = nx;
    lx_ = lx;
    n_grid_ = 1;
    for (size_t d = 0; d < nx.size(); d++) {
        dx_.push_back(lx[d] / nx[d]);
        n_grid_ *= nx[d];
    }
    if (mask) {
        mask_.resize(n_grid_);
        for (int i = 0; i < n_grid_; i++) {
            if (std::abs(mask[i]) < 1e-7) mask_[i] = 0.0;
            else if (std::abs(mask[i] - 1.0) < 1e-7) mask_[i] = 1.0;
            else LFTS_THROW("mask[" + std::to_string(i) + "] must be 0.0 or 1.0");
        }
    }
    volume_ = 1.0;
    for (double l : lx) volume_ *= l;
    const size_t dim = nx.size();
    if (bc.size() != 0 && bc.size() != 2 * dim)
        LFTS_THROW("We expect 0 or " + std::to_string(2 * dim) + " boundary conditions, but we get " +
                   std::to_string(bc.size()) + ". For each dimension, two boundary conditions (top and bottom) are required.");
    for (auto& b : bc) {
        std::transform(b.begin(), b.end(), b.begin(), [](unsigned char c) { return std::tolower(c); });
        if (b != "periodic" && b != "reflecting" && b != "absorbing")
            LFTS_THROW(b + " is an invalid boundary condition. Choose among ['periodic', 'reflecting', 'absorbing']");
    }
    for (size_t i = 0; i < bc.size(); i += 2)
        if ((bc[i] == "periodic") != (bc[i + 1] == "periodic"))
            LFTS_THROW(bc[i] + " and " + bc[i + 1] + " are an invalid boundary condition combination. If one side "
                       "imposes a periodic boundary condition, the other side must also be a periodic boundary condition.");
    if (bc.empty()) bc.assign(2 * dim, "periodic");
    bc_ = bc;
    check_fts(fts_grid_create(hip_context(), (int)dim, nx_.data(), lx_.data(), get_mask(), &grid_));
}

ComputationBox::~ComputationBox() {
    if (grid_) fts_grid_destroy(grid_);
}

int ComputationBox::get_nx(int i) const {
    if (i < 0 || i >= get_dim()) LFTS_THROW("'" + std::to_string(i) + "' is out of range.");
    return nx_[i];
}
double ComputationBox::get_lx(int i) const {
    if (i < 0 || i >= get_dim()) LFTS_THROW("'" + std::to_string(i) + "' is out of range.");
    return lx_[i];
}
double ComputationBox::get_dx(int i) const {
    if (i < 0 || i >= get_dim()) LFTS_THROW("'" + std::to_string(i) + "' is out of range.");
    return dx_[i];
}
double ComputationBox::get_dv(int i) const {
    double dv = 1.0;
    for (double d : dx_) dv *= d;
    return mask_.empty() ? dv : dv * mask_.at(i);
}

bool ComputationBox::is_periodic() const {
    for (const auto& b : bc_)
        if (b != "periodic") return false;
    return true;
}

// ComputationBox.cpp:229-262 (the volume becomes the sum of dv)
void ComputationBox::set_lx(std::vector<double> new_lx) {
    if (new_lx.size() != nx_.size())
        LFTS_THROW("The sizes of new lx (" + std::to_string(new_lx.size()) + ") and dim (" +
                   std::to_string(nx_.size()) + ") must match.");
    for (double l : new_lx)
        if (l <= 0.0) LFTS_THROW("new lx must be positive numbers");
    check_fts(fts_grid_set_lx(grid_, new_lx.data()));
    lx_ = new_lx;
    for (size_t d = 0; d < nx_.size(); d++) dx_[d] = lx_[d] / nx_[d];
    check_fts(fts_grid_volume(grid_, &volume_));
}

double ComputationBox::integral(const double* g) const {
    double r;
    check_fts(fts_grid_integral(grid_, g, 0, &r));
    return r;
}

double ComputationBox::inner_product(const double* g, const double* h) const {
    double r;
    check_fts(fts_grid_inner_product(grid_, g, h, 0, &r));
    return r;
}

double ComputationBox::multi_inner_product(int n_comp, const double* g, const double* h) const {
    double s = 0.0;
    for (int n = 0; n < n_comp; n++) s += inner_product(g + (size_t)n * n_grid_, h + (size_t)n * n_grid_);
    return s;
}

void ComputationBox::zero_mean(double* g) const {
    const double s = integral(g);
    for (int i = 0; i < n_grid_; i++) g[i] -= s / volume_;
}

// ------------------------------------------------------------------ PropagatorComputation
PropagatorComputation::PropagatorComputation(ComputationBox* cb, Molecules* molecules, PropagatorAnalyzer* analyzer,
                                             int precision, int n_streams, int n_replicas, bool realspace,
                                             bool reduce_memory)
    : cb_(cb), molecules_(molecules), analyzer_(analyzer), precision_(realspace ? FTS_FP64 : precision),
      n_streams_(n_streams), n_replicas_(n_replicas), realspace_(realspace) {
    if (n_replicas < 1) LFTS_THROW("n_replicas must be positive");
    if (!cb) LFTS_THROW("ComputationBox *cb is a null pointer");
    if (!molecules) LFTS_THROW("Molecules *molecules is a null pointer");
    if (!analyzer) LFTS_THROW("PropagatorAnalyzer *propagator_analyzer is a null pointer");
    if (reduce_memory && n_replicas != 1) LFTS_THROW("the replica solver has no reduce-memory form");
    if (realspace) {
        // CpuSolverReal.cpp:16-17
        if (molecules->get_model_name() != "continuous")
            LFTS_THROW("Real-space method only support 'continuous' chain model.");
        if (n_replicas != 1) LFTS_THROW("the real-space solver has no replica form");
    } else if (!cb->is_periodic()) {
        LFTS_THROW("Currently, pseudo-spectral method only supports periodic boundary conditions");
    }
    if (analyzer->get_model_name() != molecules->get_model_name())
        LFTS_THROW("The chain model of the analyzer and of the molecules differ.");
    auto& props = analyzer->get_computation_propagators();
    if (props.empty()) LFTS_THROW("There is no propagator code. Add polymers first.");
    auto& blocks = analyzer->get_computation_blocks();
    if (blocks.empty()) LFTS_THROW("There is no block. Add polymers first.");

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
        if (!mon.count(e.monomer_type)) LFTS_THROW("monomer_type \"" + e.monomer_type + "\" is not in bond_lengths.");
        p_mon.push_back(mon[e.monomer_type]);
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
                const std::string name = propagator_code::q_input_name_from_key(key);
                auto f = std::find(qinit_names_.begin(), qinit_names_.end(), name);
                if (f == qinit_names_.end()) {
                    qinit_names_.push_back(name);
                    f = qinit_names_.end() - 1;
                }
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
    std::vector<double> vf, alpha;
    for (int p = 0; p < molecules->get_n_polymer_types(); p++) {
        vf.push_back(molecules->get_polymer(p).get_volume_fraction());
        alpha.push_back(molecules->get_polymer(p).get_alpha());
    }
    std::vector<double> svf;
    std::vector<int> smon;
    for (int s = 0; s < molecules->get_n_solvent_types(); s++) {
        const auto sv = molecules->get_solvent(s);
        if (!mon.count(std::get<1>(sv))) LFTS_THROW("monomer_type \"" + std::get<1>(sv) + "\" is not in bond_lengths.");
        svf.push_back(std::get<0>(sv));
        smon.push_back(mon.at(std::get<1>(sv)));
    }
    Scheduler sc(props, n_streams_);
    schedule_ = sc.get_schedule();
    std::vector<int> s_off{0}, j_prop, j_from, j_to;
    for (const auto& slice : schedule_) {
        for (const auto& j : slice) {
            j_prop.push_back(prop_index_.at(std::get<0>(j)));
            j_from.push_back(std::get<1>(j));
            j_to.push_back(std::get<2>(j));
        }
        s_off.push_back((int)j_prop.size());
    }
    auto ip = [](std::vector<int>& v) {
        if (v.empty()) v.push_back(0);
        return v.data();
    };
    auto dp = [](std::vector<double>& v) {
        if (v.empty()) v.push_back(0.0);
        return v.data();
    };
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
    pl.n_slices = (int)schedule_.size();
    pl.slice_off = ip(s_off);
    pl.job_prop = ip(j_prop);
    pl.job_from = ip(j_from);
    pl.job_to = ip(j_to);
    std::vector<int> bc;
    if (realspace_)
        for (const auto& b : cb->get_boundary_conditions())
            bc.push_back(b == "periodic" ? FTS_BC_PERIODIC : b == "reflecting" ? FTS_BC_REFLECTING : FTS_BC_ABSORBING);
    if (reduce_memory) {
        check_fts(fts_solver_create_reduce_memory(cb->handle(), &pl, precision_, realspace_ ? bc.data() : nullptr,
                                                  &solver_));
    } else if (realspace_) {
        check_fts(fts_solver_create_realspace(cb->handle(), &pl, bc.data(), &solver_));
    } else if (n_replicas_ == 1) check_fts(fts_solver_create(cb->handle(), &pl, precision_, &solver_));
    else check_fts(fts_solver_create_replicas(cb->handle(), &pl, precision_, n_replicas_, &solver_));
}

std::pair<size_t, size_t> PropagatorComputation::memory_usage() const {
    size_t d = 0, h = 0;
    check_fts(fts_solver_get_memory_usage(solver_, &d, &h));
    return {d, h};
}

PropagatorComputation::~PropagatorComputation() {
    if (solver_) fts_solver_destroy(solver_);
}

void PropagatorComputation::update_laplacian_operator() { check_fts(fts_solver_update_laplacian(solver_)); }

void PropagatorComputation::compute_propagators(const std::map<std::string, const double*>& w_input,
                                                const std::map<std::string, const double*>& q_init,
                                                bool on_device) {
    for (const auto& it : analyzer_->get_computation_propagators())
        if (!w_input.count(it.second.monomer_type))
            LFTS_THROW("monomer_type \"" + it.second.monomer_type + "\" is not in w_input.");
    for (const auto& it : w_input)
        if (!molecules_->get_bond_lengths().count(it.first))
            LFTS_THROW("monomer_type \"" + it.first + "\" is not in exp_dw.");
    for (int s = 0; s < molecules_->get_n_solvent_types(); s++)
        if (!w_input.count(std::get<1>(molecules_->get_solvent(s))))
            LFTS_THROW("monomer_type \"" + std::get<1>(molecules_->get_solvent(s)) + "\" is not in w_input.");
    static std::vector<double> zeros;
    const int M = cb_->get_n_grid();
    const double* zero_dev = nullptr;
    std::vector<const double*> w;
    for (const auto& m : monomers_) {
        auto it = w_input.find(m);
        if (it != w_input.end()) {
            w.push_back(it->second);
        } else {
            if (on_device) {
                // a monomer that no propagator or solvent uses: any field will do
                zero_dev = w_input.begin()->second;
                w.push_back(zero_dev);
            } else {
                if ((int)zeros.size() < M) zeros.assign(M, 0.0);
                w.push_back(zeros.data());
            }
        }
    }
    last_w_.clear();
    if (!on_device)
        for (const auto& it : w_input) last_w_[it.first].assign(it.second, it.second + M);
    std::vector<const double*> q;
    for (const auto& name : qinit_names_) {
        auto it = q_init.find(name);
        if (it == q_init.end()) LFTS_THROW("Could not find q_init[\"" + name + "\"].");
        q.push_back(it->second);
    }
    check_fts(fts_solver_compute_propagators(solver_, w.data(), q.empty() ? nullptr : q.data(), on_device ? 1 : 0));
}

void PropagatorComputation::compute_concentrations() { check_fts(fts_solver_compute_concentrations(solver_)); }

void PropagatorComputation::compute_stress() { check_fts(fts_solver_compute_stress(solver_)); }

double PropagatorComputation::get_total_partition(int polymer) {
    if (polymer < 0 || polymer >= molecules_->get_n_polymer_types())
        LFTS_THROW("Index (" + std::to_string(polymer) + ") must be in range [0, " +
                   std::to_string(molecules_->get_n_polymer_types() - 1) + "]");
    double q;
    check_fts(fts_solver_get_partition(solver_, polymer, &q));
    return q;
}

void PropagatorComputation::get_chain_propagator(double* q_out, int polymer, int v, int u, int n) {
    Polymer& pc = molecules_->get_polymer(polymer);
    const std::string key = pc.get_propagator_key(v, u);
    auto it = prop_index_.find(key);
    if (it == prop_index_.end())
        LFTS_THROW("Could not find the propagator code '" + key +
                   "'. Disable 'aggregation' option to obtain propagator_analyzer.");
    const int N = analyzer_->get_computation_propagator(key).max_n_segment;
    const int lo = molecules_->get_model_name() == "continuous" ? 0 : 1;
    if (n < lo || n > N)
        LFTS_THROW("n (" + std::to_string(n) + ") must be in range [" + std::to_string(lo) + ", " + std::to_string(N) + "]");
    check_fts(fts_solver_get_propagator(solver_, it->second, n, 0, q_out, 0));
}

// sum of the block concentrations of the given monomer type (CpuComputationContinuous.cpp:477-545)
void PropagatorComputation::sum_blocks(int polymer, const std::string& monomer, bool with_solvents, double scale,
                                       double* out, bool on_device) {
    std::vector<int> kinds, idx;
    std::vector<double> sc;
    auto& blocks = analyzer_->get_computation_blocks();
    for (size_t b = 0; b < block_keys_.size(); b++) {
        const BlockKey& k = block_keys_[b];
        if (polymer >= 0 && std::get<0>(k) != polymer) continue;
        if (propagator_code::monomer_type_from_key(std::get<1>(k)) == monomer && blocks.at(k).n_segment_right != 0) {
            kinds.push_back(FTS_SRC_BLOCK);
            idx.push_back((int)b);
            sc.push_back(scale);
        }
    }
    if (with_solvents)
        for (int s = 0; s < molecules_->get_n_solvent_types(); s++)
            if (std::get<1>(molecules_->get_solvent(s)) == monomer) {
                kinds.push_back(FTS_SRC_SOLVENT);
                idx.push_back(s);
                sc.push_back(1.0);
            }
    check_fts(fts_solver_sum_concentrations(solver_, (int)kinds.size(), kinds.data(), idx.data(), sc.data(), out,
                                            on_device ? 1 : 0));
}

void PropagatorComputation::concentration_terms(const std::string& monomer, double scale, std::vector<int>& kinds,
                                                std::vector<int>& idx, std::vector<double>& sc) const {
    auto& blocks = analyzer_->get_computation_blocks();
    for (size_t b = 0; b < block_keys_.size(); b++) {
        const BlockKey& k = block_keys_[b];
        if (propagator_code::monomer_type_from_key(std::get<1>(k)) == monomer && blocks.at(k).n_segment_right != 0) {
            kinds.push_back(FTS_SRC_BLOCK);
            idx.push_back((int)b);
            sc.push_back(scale);
        }
    }
    for (int s = 0; s < molecules_->get_n_solvent_types(); s++)
        if (std::get<1>(molecules_->get_solvent(s)) == monomer) {
            kinds.push_back(FTS_SRC_SOLVENT);
            idx.push_back(s);
            sc.push_back(scale);
        }
}

void PropagatorComputation::get_total_concentration(const std::string& monomer_type, double* phi, bool on_device) {
    sum_blocks(-1, monomer_type, true, 1.0, phi, on_device);
}

void PropagatorComputation::get_total_concentration(int polymer, const std::string& monomer_type, double* phi,
                                                    bool on_device) {
    const int P = molecules_->get_n_polymer_types();
    if (polymer < 0 || polymer > P - 1)
        LFTS_THROW("Index (" + std::to_string(polymer) + ") must be in range [0, " + std::to_string(P - 1) + "]");
    sum_blocks(polymer, monomer_type, false, 1.0, phi, on_device);
}

// CpuComputationContinuous.cpp:546-577
void PropagatorComputation::get_total_concentration_gce(double fugacity, int polymer, const std::string& monomer_type,
                                                        double* phi) {
    const int P = molecules_->get_n_polymer_types();
    if (polymer < 0 || polymer > P - 1)
        LFTS_THROW("Index (" + std::to_string(polymer) + ") must be in range [0, " + std::to_string(P - 1) + "]");
    Polymer& pc = molecules_->get_polymer(polymer);
    const double norm = fugacity / pc.get_volume_fraction() * pc.get_alpha() * get_total_partition(polymer);
    sum_blocks(polymer, monomer_type, false, norm, phi, false);
}

// CpuComputationContinuous.cpp:578-612: one array per block of the polymer, in block order
void PropagatorComputation::get_block_concentration(int polymer, double* phi) {
    const int P = molecules_->get_n_polymer_types();
    if (polymer < 0 || polymer > P - 1)
        LFTS_THROW("Index (" + std::to_string(polymer) + ") must be in range [0, " + std::to_string(P - 1) + "]");
    if (analyzer_->is_aggregated())
        LFTS_THROW("Disable 'aggregation' option to invoke 'get_block_concentration'.");
    Polymer& pc = molecules_->get_polymer(polymer);
    const int M = cb_->get_n_grid();
    const auto& blocks = pc.get_blocks();
    for (size_t b = 0; b < blocks.size(); b++) {
        std::string kl = pc.get_propagator_key(blocks[b].v, blocks[b].u);
        std::string kr = pc.get_propagator_key(blocks[b].u, blocks[b].v);
        if (kl < kr) kl.swap(kr);
        const auto it = std::find(block_keys_.begin(), block_keys_.end(), std::make_tuple(polymer, kl, kr));
        if (it == block_keys_.end()) LFTS_THROW("block (" + kl + ", " + kr + ") not found");
        const int kind = FTS_SRC_BLOCK, idx = (int)(it - block_keys_.begin());
        const double one = 1.0;
        check_fts(fts_solver_sum_concentrations(solver_, 1, &kind, &idx, &one, phi + b * (size_t)M, 0));
    }
}

double PropagatorComputation::get_solvent_partition(int s) {
    double q;
    check_fts(fts_solver_get_solvent_partition(solver_, s, &q));
    return q;
}

void PropagatorComputation::get_solvent_concentration(int s, double* phi) {
    const int S = molecules_->get_n_solvent_types();
    if (s < 0 || s > S - 1)
        LFTS_THROW("Index (" + std::to_string(s) + ") must be in range [0, " + std::to_string(S - 1) + "]");
    const int kind = FTS_SRC_SOLVENT;
    const double one = 1.0;
    check_fts(fts_solver_sum_concentrations(solver_, 1, &kind, &s, &one, phi, 0));
}

std::vector<double> PropagatorComputation::dq_dl(int p) {
    std::vector<double> out(3, 0.0);
    check_fts(fts_solver_get_dq_dl(solver_, p, out.data()));
    out.resize(cb_->get_dim());
    return out;
}

// PropagatorComputation.cpp:35-52
std::vector<double> PropagatorComputation::get_stress() {
    std::vector<double> stress(cb_->get_dim(), 0.0);
    for (int p = 0; p < molecules_->get_n_polymer_types(); p++) {
        Polymer& pc = molecules_->get_polymer(p);
        const std::vector<double> d = dq_dl(p);
        const double Q = get_total_partition(p);
        for (int k = 0; k < cb_->get_dim(); k++) stress[k] += d[k] * pc.get_volume_fraction() / pc.get_alpha() / Q;
    }
    return stress;
}

// PropagatorComputation.cpp:53-69
std::vector<double> PropagatorComputation::get_stress_gce(const std::vector<double>& fugacities) {
    std::vector<double> stress(cb_->get_dim(), 0.0);
    if ((int)fugacities.size() != molecules_->get_n_polymer_types())
        LFTS_THROW("The number of fugacities must match the number of polymer types.");
    for (int p = 0; p < molecules_->get_n_polymer_types(); p++) {
        const std::vector<double> d = dq_dl(p);
        for (int k = 0; k < cb_->get_dim(); k++) stress[k] += fugacities[p] * d[k];
    }
    return stress;
}

// CpuComputationContinuous.cpp:755-817, CpuComputationDiscrete.cpp:1027-1089: Q from every
// contour position of every block must agree within 1e-7
bool PropagatorComputation::check_total_partition() {
    const int M = cb_->get_n_grid();
    const bool continuous = molecules_->get_model_name() == "continuous";
    std::vector<std::vector<double>> totals(molecules_->get_n_polymer_types());
    std::vector<double> a(M), b(M), inv_e;
    auto& blocks = analyzer_->get_computation_blocks();
    for (size_t i = 0; i < block_keys_.size(); i++) {
        const BlockKey& k = block_keys_[i];
        const ComputationBlock& blk = blocks.at(k);
        const int L = prop_index_.at(std::get<1>(k)), R = prop_index_.at(std::get<2>(k));
        const int nl = blk.n_segment_left, nr = blk.n_segment_right;
        const int n_prop = (int)blk.v_u.size();
        if (!continuous) {
            // discrete weight 1/E = exp(+w ds) of the block's monomer
            auto w = last_w_.find(blk.monomer_type);
            if (w == last_w_.end())
                LFTS_THROW("check_total_partition needs the fields of the last compute_propagators call on the host.");
            inv_e.resize(M);
            for (int i = 0; i < M; i++) inv_e[i] = std::exp(w->second[i] * molecules_->get_ds());
        }
        for (int n = continuous ? 0 : 1; n <= nr; n++) {
            check_fts(fts_solver_get_propagator(solver_, L, continuous ? nl - n : nl - n + 1, 0, a.data(), 0));
            check_fts(fts_solver_get_propagator(solver_, R, n, 0, b.data(), 0));
            if (!continuous)
                for (int i = 0; i < M; i++) b[i] *= inv_e[i];
            const double s = cb_->inner_product(a.data(), b.data());
            totals[std::get<0>(k)].push_back(s * blk.n_repeated / cb_->get_volume() / n_prop);
        }
    }
    std::cout << "Polymer id: maximum,  minimum, and difference of total partitions" << std::endl;
    bool ok = true;
    for (size_t p = 0; p < totals.size(); p++) {
        double mx = -1e20, mn = 1e20;
        for (double t : totals[p]) {
            mx = std::max(mx, t);
            mn = std::min(mn, t);
        }
        std::cout << "\t" << p << ": " << mx << ", " << mn << ", " << std::abs(mx - mn) << std::endl;
        if (std::abs(mx - mn) > 1e-7) ok = false;
    }
    return ok;
}

// ------------------------------------------------------------------ AndersonMixing
AndersonMixing::AndersonMixing(int n_var, int max_hist, double start_error, double mix_min, double mix_init,
                               bool reduce_memory)
    : n_var_(n_var) {
    if (reduce_memory)
        check_fts(fts_am_create_reduce_memory(hip_context(), n_var, max_hist, start_error, mix_min, mix_init, &am_));
    else
        check_fts(fts_am_create(hip_context(), n_var, max_hist, start_error, mix_min, mix_init, &am_));
}

AndersonMixing::~AndersonMixing() {
    if (am_) fts_am_destroy(am_);
}

void AndersonMixing::reset_count() { check_fts(fts_am_reset(am_)); }

void AndersonMixing::calculate_new_fields(double* w_new, const double* w_current, const double* w_deriv,
                                          double old_error_level, double error_level, bool on_device) {
    check_fts(fts_am_calculate_new_fields(am_, w_new, w_current, w_deriv, old_error_level, error_level,
                                          on_device ? 1 : 0));
}

// ------------------------------------------------------------------ L-FTS saddle point
DeviceSaddlePoint::DeviceSaddlePoint(PropagatorComputation* solver, AndersonMixing* am, const FieldTheoryTables& t,
                                     const std::vector<double>& am_params) {
    if (!solver) LFTS_THROW("solver is required");
    NR_ = solver->n_replicas();
    n_blocks_ = solver->n_blocks_per_replica();
    if (NR_ == 1 && !am) LFTS_THROW("am is required");
    if (NR_ > 1 && am_params.size() != 4)
        LFTS_THROW("a replica solver needs am_params = [max_hist, start_error, mix_min, mix_init]");
    S_ = (int)t.monomer_types.size();
    R_ = (int)t.real_idx.size();
    I_ = (int)t.imag_idx.size();
    M_ = solver->get_n_grid();
    Molecules* mol = solver->molecules();
    P_ = mol->get_n_polymer_types();
    auto need = [&](const std::vector<double>& v, size_t n, const char* what) {
        if (v.size() != n) LFTS_THROW(std::string(what) + " has size " + std::to_string(v.size()) + ", expected " +
                                      std::to_string(n));
    };
    if (S_ < 1 || S_ > FTS_MAX_TYPES) LFTS_THROW("the number of monomer types must be in [1, 8]");
    need(t.matrix_a, (size_t)S_ * S_, "matrix_a");
    need(t.real_cw, (size_t)R_ * S_, "real_cw");
    need(t.real_cp, (size_t)R_ * S_, "real_cp");
    need(t.real_c0, (size_t)R_, "real_c0");
    need(t.imag_cw, (size_t)I_ * S_, "imag_cw");
    need(t.imag_cp, (size_t)I_ * S_, "imag_cp");
    need(t.imag_c0, (size_t)I_, "imag_c0");
    need(t.dt_scaling, (size_t)S_, "dt_scaling");
    need(t.h_lin, (size_t)S_, "h_lin");
    need(t.h_sq, (size_t)S_, "h_sq");
    if (NR_ == 1 && I_ * M_ != am->get_n_var())
        LFTS_THROW("Anderson mixing has n_var " + std::to_string(am->get_n_var()) + ", expected " +
                   std::to_string(I_ * M_));
    std::map<std::string, int> type_index;
    for (int i = 0; i < S_; i++) type_index[t.monomer_types[i]] = i;
    // solver monomer fields as rows over w_aux: w = A w_aux for the S types, random
    // copolymers mix the rows of their constituent types (lfts.py:534-545)
    const auto& sm = solver->monomer_types();
    std::vector<double> w_map(sm.size() * S_, 0.0);
    for (size_t m = 0; m < sm.size(); m++) {
        auto add_row = [&](int type, double f) {
            for (int j = 0; j < S_; j++) w_map[m * S_ + j] += f * t.matrix_a[(size_t)type * S_ + j];
        };
        auto ti = type_index.find(sm[m]);
        auto ri = t.random_fraction.find(sm[m]);
        if (ti != type_index.end()) {
            add_row(ti->second, 1.0);
        } else if (ri != t.random_fraction.end()) {
            for (const auto& f : ri->second) {
                auto fi = type_index.find(f.first);
                if (fi == type_index.end()) LFTS_THROW("random copolymer type \"" + f.first + "\" is unknown");
                add_row(fi->second, f.second);
            }
        } else {
            LFTS_THROW("monomer type \"" + sm[m] + "\" is neither a field type nor a random copolymer");
        }
    }
    // concentration of each type: its own blocks and solvents, plus the random
    // copolymer concentrations times their fractions (lfts.py:550-565)
    std::vector<int> off{0}, kinds, idx;
    std::vector<double> sc;
    for (int i = 0; i < S_; i++) {
        solver->concentration_terms(t.monomer_types[i], 1.0, kinds, idx, sc);
        for (const auto& r : t.random_fraction) {
            auto f = r.second.find(t.monomer_types[i]);
            if (f != r.second.end()) solver->concentration_terms(r.first, f->second, kinds, idx, sc);
        }
        off.push_back((int)kinds.size());
    }
    std::vector<double> weight(P_);
    for (int p = 0; p < P_; p++)
        weight[p] = mol->get_polymer(p).get_volume_fraction() / mol->get_polymer(p).get_alpha();

    fts_field_theory ft{};
    ft.n_types = S_;
    ft.n_real = R_;
    ft.n_imag = I_;
    ft.real_idx = t.real_idx.data();
    ft.imag_idx = t.imag_idx.data();
    ft.n_solver_fields = (int)sm.size();
    ft.w_map = w_map.data();
    ft.phi_term_off = off.data();
    ft.phi_term_kind = kinds.data();
    ft.phi_term_index = idx.data();
    ft.phi_term_scale = sc.data();
    ft.real_cw = t.real_cw.data();
    ft.real_cp = t.real_cp.data();
    ft.real_c0 = t.real_c0.data();
    ft.imag_cw = t.imag_cw.data();
    ft.imag_cp = t.imag_cp.data();
    ft.imag_c0 = t.imag_c0.data();
    ft.dt_scaling = t.dt_scaling.data();
    ft.h_lin = t.h_lin.data();
    ft.h_sq = t.h_sq.data();
    ft.h_const = t.h_const;
    ft.n_polymers = P_;
    ft.polymer_weight = weight.data();
    if (NR_ == 1)
        check_fts(fts_saddle_create(hip_context(), solver->handle(), am->handle(), M_, &ft, &saddle_));
    else
        check_fts(fts_saddle_create_replicas(hip_context(), solver->handle(), NR_, solver->n_blocks_per_replica(),
                                             mol->get_n_solvent_types(), (int)am_params[0], am_params[1],
                                             am_params[2], am_params[3], M_, &ft, &saddle_));
}

void DeviceSaddlePoint::restore_replicas(const int* flags) { check_fts(fts_saddle_restore_replicas(saddle_, flags)); }

DeviceSaddlePoint::~DeviceSaddlePoint() {
    if (saddle_) fts_saddle_destroy(saddle_);
}

void DeviceSaddlePoint::set_fields(const double* w_aux) { check_fts(fts_saddle_set_fields(saddle_, w_aux, 0)); }
void DeviceSaddlePoint::get_fields(double* w_aux) { check_fts(fts_saddle_get_fields(saddle_, w_aux, 0)); }
void DeviceSaddlePoint::get_concentration(int type, double* phi) {
    check_fts(fts_saddle_get_concentration(saddle_, type, phi, 0));
}
void DeviceSaddlePoint::set_noise_prev(const double* noise) { check_fts(fts_saddle_set_noise(saddle_, noise, 0)); }
void DeviceSaddlePoint::get_noise_prev(double* noise) { check_fts(fts_saddle_get_noise(saddle_, noise, 0)); }
void DeviceSaddlePoint::backup() { check_fts(fts_saddle_backup(saddle_)); }
void DeviceSaddlePoint::restore() { check_fts(fts_saddle_restore(saddle_)); }
void DeviceSaddlePoint::find(double tolerance, int max_iter, int verbose, double* log, int* n_iter, double* error,
                             double* ham) {
    check_fts(fts_saddle_find(saddle_, tolerance, max_iter, verbose, log, n_iter, error, ham));
}
void DeviceSaddlePoint::langevin_update(const double* noise, double dt) {
    check_fts(fts_saddle_langevin_update(saddle_, noise, 0, dt));
}
void DeviceSaddlePoint::field_moments(double* mean, double* mean_sq) {
    check_fts(fts_saddle_field_moments(saddle_, mean, mean_sq));
}

void DeviceSaddlePoint::set_low_precision_solver(PropagatorComputation* lo, double switch_error) {
    if (lo && lo->n_replicas() != NR_) LFTS_THROW("the low-precision solver must have the same number of replicas");
    if (lo && lo->n_blocks_per_replica() != n_blocks_)
        LFTS_THROW("the low-precision solver must be built from the same molecules and analyzer");
    check_fts(fts_saddle_set_low_precision(saddle_, lo ? lo->handle() : nullptr, switch_error));
}

void DeviceSaddlePoint::set_graphs(bool on) { check_fts(fts_saddle_set_graphs(saddle_, on ? 1 : 0)); }
void DeviceSaddlePoint::prepare_guess(LearnedGuess& g) { check_fts(fts_saddle_prepare_guess(saddle_, g.handle())); }
void DeviceSaddlePoint::predict_guess(LearnedGuess& g, int mode) {
    check_fts(fts_saddle_predict_guess(saddle_, g.handle(), mode));
}
void DeviceSaddlePoint::set_residual_guess(LearnedGuess* g, bool record) {
    check_fts(fts_saddle_set_residual_guess(saddle_, g ? g->handle() : nullptr, record ? 1 : 0));
}
std::vector<double> DeviceSaddlePoint::guess_sample() {
    std::vector<double> out(4 * (size_t)M_);
    check_fts(fts_saddle_get_guess_sample(saddle_, out.data()));
    return out;
}
double DeviceSaddlePoint::train_guess(LearnedGuess& g, double lr, int iters, int replay) {
    double loss = 0.0;
    check_fts(fts_saddle_train_guess(saddle_, g.handle(), lr, iters, replay, &loss));
    return loss;
}

// ------------------------------------------------------------------ learned guess
LearnedGuess::LearnedGuess(std::vector<int> nx, std::vector<double> in_scales, double out_scale) {
    if (nx.size() != 3) LFTS_THROW("the learned guess needs a 3D grid");
    if (in_scales.empty()) LFTS_THROW("the learned guess needs at least one input scale");
    M_ = nx[0] * nx[1] * nx[2];
    n_in_ = (int)in_scales.size();
    check_fts(fts_cnn_create_inputs(hip_context(), nx.data(), n_in_, in_scales.data(), out_scale, &cnn_));
    check_fts(fts_cnn_param_count(cnn_, &n_params_));
}
LearnedGuess::~LearnedGuess() {
    if (cnn_) fts_cnn_destroy(cnn_);
}
void LearnedGuess::set_params(const float* p) { check_fts(fts_cnn_set_params(cnn_, p)); }
void LearnedGuess::get_params(float* p) { check_fts(fts_cnn_get_params(cnn_, p)); }
void LearnedGuess::predict(const double* w_minus, double* w_plus) {
    check_fts(fts_cnn_predict(cnn_, w_minus, w_plus, 0));
}
double LearnedGuess::gradients(const double* w_minus, const double* target, float* grads) {
    double loss = 0.0;
    check_fts(fts_cnn_gradients(cnn_, w_minus, target, 0, &loss, grads));
    return loss;
}
double LearnedGuess::train_step(const double* w_minus, const double* target, double lr) {
    double loss = 0.0;
    check_fts(fts_cnn_train_step(cnn_, w_minus, target, 0, lr, &loss));
    return loss;
}

// ------------------------------------------------------------------ factory
ComputationBox* AbstractFactory::create_computation_box(std::vector<int> nx, std::vector<double> lx,
                                                        std::vector<std::string> bc, const double* mask) {
    return new ComputationBox(nx, lx, bc, mask);
}

Molecules* AbstractFactory::create_molecules_information(const std::string& chain_model, double ds,
                                                         const std::map<std::string, double>& bond_lengths) {
    return new Molecules(chain_model, ds, bond_lengths);
}

PropagatorAnalyzer* AbstractFactory::create_propagator_analyzer(Molecules* molecules, bool aggregate) {
    return new PropagatorAnalyzer(molecules, aggregate);
}

// propagators run on at most LFTS_HIP_LANES concurrent lanes per schedule slice
static int lanes() {
    const char* env = std::getenv("LFTS_HIP_LANES");
    const int n = env ? std::atoi(env) : 8;
    return n >= 1 ? n : 8;
}

PropagatorComputation* AbstractFactory::create_replica_solver(ComputationBox* cb, Molecules* molecules,
                                                              PropagatorAnalyzer* analyzer, int n_replicas) {
    return new PropagatorComputation(cb, molecules, analyzer, precision_, lanes(), n_replicas);
}

PropagatorComputation* AbstractFactory::create_pseudospectral_solver(ComputationBox* cb, Molecules* molecules,
                                                                     PropagatorAnalyzer* analyzer) {
    return new PropagatorComputation(cb, molecules, analyzer, precision_, lanes(), 1, false, reduce_memory_usage_);
}

// CpuComputationContinuous(cb, molecules, analyzer, "realspace") (CpuComputationContinuous.cpp:25-26)
PropagatorComputation* AbstractFactory::create_realspace_solver(ComputationBox* cb, Molecules* molecules,
                                                                PropagatorAnalyzer* analyzer) {
    return new PropagatorComputation(cb, molecules, analyzer, FTS_FP64, lanes(), 1, true, reduce_memory_usage_);
}

AndersonMixing* AbstractFactory::create_anderson_mixing(int n_var, int max_hist, double start_error, double mix_min,
                                                        double mix_init) {
    return new AndersonMixing(n_var, max_hist, start_error, mix_min, mix_init, reduce_memory_usage_);
}

void AbstractFactory::display_info() const {
    char info[256] = {0};
    fts_context_info(hip_context(), info, sizeof(info));
    std::cout << "========== HIP platform (" << name_ << ") ==========" << std::endl;
    std::cout << "Device: " << info << std::endl;
    std::cout << "Propagators and FFTs in " << (precision_ == FTS_FP64 ? "FP64" : "FP32")
              << "; Q, concentrations, stress and Anderson mixing in FP64" << std::endl;
    std::cout << "Concurrent propagator lanes per schedule slice: " << lanes() << std::endl;
    if (reduce_memory_usage_)
        std::cout << "Reduce memory usage: propagators and Anderson-mixing history in pinned host memory" << std::endl;
}

std::vector<std::string> PlatformSelector::avail_platforms() { return {"hip", "hip-fp32"}; }

AbstractFactory* PlatformSelector::create_factory(const std::string& platform) { return create_factory(platform, false); }

AbstractFactory* PlatformSelector::create_factory(const std::string& platform, bool reduce_memory_usage) {
    if (platform == "hip") return new AbstractFactory(platform, FTS_FP64, reduce_memory_usage);
    if (platform == "hip-fp32") return new AbstractFactory(platform, FTS_FP32, reduce_memory_usage);
    LFTS_THROW("Could not find platform '" + platform + "'");
}

}  // namespace lfts
