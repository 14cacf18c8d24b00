// Propagator computation on the device: the plan (propagators, blocks, schedule)
// is compiled once into a list of batched launches; compute_propagators replays it.
//
// Follows CpuComputationContinuous.cpp:148-726 and CpuComputationDiscrete.cpp:192-996
// for what is computed; the batching is this file's own: all jobs of one schedule
// slice (Scheduler.cpp:137-188) advance together, one launch per pass for the
// whole slice, and consecutive steps of a propagator share a slab pass.
// With boundary conditions given, continuous chains take real-space
// Crank-Nicolson steps instead (CpuComputationContinuous.cpp:25-26).
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <map>

#include "fts_internal.hpp"
#include "fts_real.hpp"
#include "fts_reduce.hpp"

namespace fts {

namespace {

template <typename T>
class Solver final : public SolverBase {
    using T2 = typename vec2<T>::type;

public:
    Solver(Grid* grid, const fts_plan* pl, const int* bc, bool reduce_memory);
    void update_laplacian() override {
        if (real_) build_real_tables();
        else bt_.build(*g_, bsq_, ds_);
        tables_version_++;
    }
    bool graphs_on() const override { return graphs_; }
    void results_changed() override { Q_valid_ = false; }
    std::vector<double> launch_signature() const override {
        const GraphKey k = graph_key(nullptr, nullptr);
        return {k.volume, (double)k.bd, (double)k.nl, (double)k.bd_fwd, (double)k.bd_inv2, (double)tables_version_};
    }
    void compute_propagators(const double* const* w, const double* const* q_init, bool on_device) override;
    void compute_concentrations() override;
    void compute_stress() override;
    double partition(int p) override;
    double solvent_partition(int s) override;
    void dq_dl(int p, double* out) override;
    void sum_concentrations(int n, const int* kind, const int* idx, const double* scale, double* out,
                            bool on_device) override;
    void sum_concentrations_replicas(int n, const int* kind, const int* idx, const double* scale, double* out,
                                     int n_rep, int block_step, int solvent_step, size_t out_step) override;
    void propagator(int prop, int n, int half, double* out, bool on_device) override;
    void set_graphs(bool on) override {
        graphs_ = on;
        if (!on) drop_graphs();
    }
    ~Solver() override { drop_graphs(); }
    void memory_usage(size_t* device, size_t* host) const override {
        size_t d = 0, h = 0;
        for (const DeviceBuffer* b : {&qstore_, &hstore_, &w64_, &w32_, &ew_, &qinit_, &ws_, &phib_, &phis_, &Qd_,
                                      &stress_out_, &rtab_})
            (b->pinned_host() ? h : d) += b->bytes();
        *device = d;
        *host = h;
    }

private:
    struct Prop {
        int monomer, max_n, kind, qinit;
        std::vector<std::array<int, 3>> deps;
        std::vector<int> je;
    };
    struct Block {
        int polymer, left, right, nl, nr, rep, monomer, nvu;
    };
    struct STerm {
        int x1p, x1n, x1h, x2p, x2n, x2h;  // (prop, n, half?) of both slices
        double bsq, weight;
        int table, block;
    };
    enum { L_INIT, L_SLAB_FWD, L_SLAB_INV1, L_SLAB_INV2, L_XPASS, L_REAL };
    struct Launch {
        int type, idx;
    };

    T* q(int p, int n) const { return qstore_.as<T>() + qoff_[p] + (size_t)n * M_; }
    T* h(int p, int n) const {
        auto it = hoff_.find({p, n});
        FTS_REQUIRE(it != hoff_.end(), "half-bond slice not allocated");
        return hstore_.as<T>() + it->second;
    }
    T* slice_ptr(int p, int n, int half) const { return half ? h(p, n) : q(p, n); }
    const T* w(int m) const { return wT_ + (size_t)m * M_; }
    // discrete FP64 chains: exp(-w ds) of monomer m, made once per compute_propagators (the FP32
    // kernels compute their single-precision exponential on the fly: a table measures the same there)
    bool table() const { return model_ == FTS_DISCRETE && std::is_same<T, double>::value; }
    const T* ew(int m) const { return table() ? ew_.as<T>() + (size_t)m * M_ : w(m); }
    T ew_coef() const { return table() ? (T)-1 : (T)ds_; }   // negative: the array is the factor
    T2* slot(int s) const { return ws_.as<T2>() + (size_t)s * Ms_; }
    const T* mask() const { return g_->mask_h.empty() ? nullptr : g_->mask<T>(); }
    int full(int m) const { return 2 * m; }
    int half(int m) const { return 2 * m + 1; }

    void emit_slab(int mode, const std::vector<SlabItem<T>>& items);
    void emit_x(const std::vector<XItem<T>>& items);
    void emit_init(const std::vector<InitItem<T>>& items);
    void emit_real(const std::vector<RealItem>& items);
    void build_real_tables();
    void launch_real(const RealBatch& b);
    void build_slice(const std::vector<std::array<int, 3>>& jobs);
    void build_stress_terms();
    void run_launches();
    void issue_propagators(const double* const* w, const double* const* q_init, bool on_device);
    void issue_concentrations();

    // HIP graphs of the two evaluation calls, valid while their key holds: the input
    // addresses, the box volume (a by-value kernel argument) and the FFT launch shapes
    struct GraphKey {
        std::vector<const void*> ptrs;
        double volume = 0.0;
        int bd = 0, nl = 0, bd_fwd = 0, bd_inv2 = 0;
        bool operator==(const GraphKey& o) const {
            return ptrs == o.ptrs && volume == o.volume && bd == o.bd && nl == o.nl && bd_fwd == o.bd_fwd &&
                   bd_inv2 == o.bd_inv2;
        }
    };
    GraphKey graph_key(const double* const* w, const double* const* q_init) const;
    bool graph_ok() const;
    template <typename F>
    void replay_or_capture(hipGraphExec_t& exec, GraphKey& have, const GraphKey& want, F&& issue);
    void drop_graphs();
    bool graphs_ = false;
    hipGraphExec_t gprop_ = nullptr, gconc_ = nullptr;
    GraphKey kprop_, kconc_;

    Grid* g_;
    int model_;
    double ds_;
    int M_;
    size_t Ms_;   // spectrum workspace slot size (the grid's padded layout)
    int n_mon_, n_qinit_;
    std::vector<double> bsq_;
    std::vector<Prop> props_;
    std::vector<Block> blocks_;
    std::vector<double> vf_, alpha_, svf_;
    std::vector<int> smon_;
    std::vector<std::vector<std::array<int, 3>>> slices_;

    DeviceBuffer qstore_, hstore_, w64_, w32_, ew_, qinit_, ws_, phib_, phis_, Qd_, stress_out_;
    DeviceBuffer stress_part_;   // partial sums of the fused stress pencil pass (wave-local grids)
    T* wT_ = nullptr;
    std::vector<size_t> qoff_;
    std::map<std::pair<int, int>, size_t> hoff_;
    int n_ws_;
    BondTables bt_;

    std::vector<Launch> launches_;
    std::vector<InitBatch<T>> inits_;
    std::vector<SlabBatch<T>> slabs_;
    std::vector<XBatch<T>> xs_;
    // real-space Crank-Nicolson form
    bool real_ = false;
    bool real_plain_ = false;   // env FTS_REAL_PLAIN_SWEEPS=1: one thread per line on every axis
    std::vector<int> bc_;
    RealGeom rg_{};
    DeviceBuffer rtab_;
    std::vector<RealBatch> reals_;
    std::vector<int> part_block_;   // per polymer: block whose slices give Q, -1 none
    std::vector<STerm> sterms_;
    std::vector<double> dqdl_;
    std::vector<double> Qh_;
    bool Q_valid_ = false;
    long long tables_version_ = 0;   // bumped by update_laplacian (tables rebuilt)
};

template <typename T>
Solver<T>::Solver(Grid* grid, const fts_plan* pl, const int* bc, bool reduce_memory) : g_(grid) {
    model_ = pl->model;
    if (bc) {
        // CpuSolverReal.cpp:16-17
        FTS_REQUIRE(model_ == FTS_CONTINUOUS, "Real-space method only support 'continuous' chain model.");
        FTS_REQUIRE((std::is_same<T, double>::value), "the real-space method runs in FP64");
        real_ = true;
        real_plain_ = std::getenv("FTS_REAL_PLAIN_SWEEPS") != nullptr;
        for (int i = 0; i < 2 * g_->dim; i++) {
            FTS_REQUIRE(bc[i] >= BC_PERIODIC && bc[i] <= BC_ABSORBING, "invalid boundary condition code");
            bc_.push_back(bc[i]);
        }
        for (int a = 0; a < g_->dim; a++)
            FTS_REQUIRE((bc_[2 * a] == BC_PERIODIC) == (bc_[2 * a + 1] == BC_PERIODIC),
                        "If one side imposes a periodic boundary condition, the other side must also be a periodic "
                        "boundary condition.");
    } else {
        g_->require_fft();
    }
    ds_ = pl->ds;
    M_ = g_->M;
    Ms_ = g_->Ms;
    n_mon_ = pl->n_monomers;
    n_qinit_ = pl->n_qinit;
    FTS_REQUIRE(model_ == FTS_CONTINUOUS || model_ == FTS_DISCRETE, "invalid chain model");
    FTS_REQUIRE(n_mon_ > 0 && pl->n_props > 0 && pl->n_blocks > 0, "empty plan");
    for (int m = 0; m < n_mon_; m++) bsq_.push_back(pl->segment_length[m] * pl->segment_length[m]);
    for (int p = 0; p < pl->n_props; p++) {
        Prop pr;
        pr.monomer = pl->prop_monomer[p];
        pr.max_n = pl->prop_max_n[p];
        pr.kind = pl->prop_kind[p];
        pr.qinit = pl->prop_qinit ? pl->prop_qinit[p] : -1;
        FTS_REQUIRE(pr.monomer >= 0 && pr.monomer < n_mon_, "propagator monomer out of range");
        FTS_REQUIRE(pr.max_n >= 0, "negative max_n_segment");
        for (int d = 0; d < pl->prop_n_deps[p]; d++) {
            const int k = pl->prop_dep_off[p] + d;
            FTS_REQUIRE(pl->dep_prop[k] >= 0 && pl->dep_prop[k] < p, "dependency must precede its propagator");
            pr.deps.push_back({pl->dep_prop[k], pl->dep_n[k], pl->dep_rep[k]});
        }
        for (int d = 0; d < pl->prop_n_je[p]; d++) pr.je.push_back(pl->je[pl->prop_je_off[p] + d]);
        if (pr.kind == FTS_LEAF_QINIT) FTS_REQUIRE(pr.qinit >= 0 && pr.qinit < n_qinit_, "q_init index out of range");
        props_.push_back(pr);
    }
    for (int b = 0; b < pl->n_blocks; b++)
        blocks_.push_back({pl->blk_polymer[b], pl->blk_left[b], pl->blk_right[b], pl->blk_n_left[b],
                           pl->blk_n_right[b], pl->blk_rep[b], pl->blk_monomer[b], pl->blk_n_vu[b]});
    vf_.assign(pl->pol_vf, pl->pol_vf + pl->n_polymers);
    alpha_.assign(pl->pol_alpha, pl->pol_alpha + pl->n_polymers);
    svf_.assign(pl->sol_vf, pl->sol_vf + pl->n_solvents);
    smon_.assign(pl->sol_monomer, pl->sol_monomer + pl->n_solvents);
    size_t jmax = 1;
    for (int s = 0; s < pl->n_slices; s++) {
        std::vector<std::array<int, 3>> jobs;
        for (int j = pl->slice_off[s]; j < pl->slice_off[s + 1]; j++) jobs.push_back({pl->job_prop[j], pl->job_from[j], pl->job_to[j]});
        jmax = std::max(jmax, jobs.size());
        slices_.push_back(jobs);
    }

    // storage: q[p][0..max_n] contiguous per propagator; half-bond slices where needed
    size_t tot = 0;
    for (auto& pr : props_) {
        qoff_.push_back(tot);
        tot += (size_t)(pr.max_n + 1) * M_;
    }
    // reduce-memory form (CudaComputationReduceMemoryContinuous.cu:55-70): every propagator slice in pinned
    // host memory, which the kernels write and read in place over the host link
    if (reduce_memory) qstore_.allocate_host_pinned(sizeof(T) * tot);
    else qstore_.allocate(sizeof(T) * tot);
    size_t htot = 0;
    if (model_ == FTS_DISCRETE) {
        for (int p = 0; p < (int)props_.size(); p++) {
            if (!props_[p].deps.empty()) {
                hoff_[{p, 0}] = htot;
                htot += M_;
            }
            for (int n : props_[p].je)
                if (n >= 1 && n <= props_[p].max_n && !hoff_.count({p, n})) {
                    hoff_[{p, n}] = htot;
                    htot += M_;
                }
        }
    }
    if (htot) {
        if (reduce_memory) hstore_.allocate_host_pinned(sizeof(T) * htot);
        else hstore_.allocate(sizeof(T) * htot);
    }
    w64_.allocate(sizeof(double) * n_mon_ * (size_t)M_);
    if (std::is_same<T, double>::value) {
        wT_ = reinterpret_cast<T*>(w64_.as<double>());
    } else {
        w32_.allocate(sizeof(float) * n_mon_ * (size_t)M_);
        wT_ = reinterpret_cast<T*>(w32_.as<float>());
    }
    if (table()) ew_.allocate(sizeof(T) * n_mon_ * (size_t)M_);
    if (n_qinit_) qinit_.allocate(sizeof(double) * n_qinit_ * (size_t)M_);
    n_ws_ = std::max<int>(2 * (int)jmax, 16);
    ws_.allocate(sizeof(T2) * (size_t)n_ws_ * Ms_);
    phib_.allocate(sizeof(double) * blocks_.size() * (size_t)M_);
    if (!svf_.empty()) phis_.allocate(sizeof(double) * svf_.size() * (size_t)M_);
    Qd_.allocate(sizeof(double) * (vf_.size() + svf_.size() + 1));
    Qh_.assign(vf_.size() + svf_.size(), 0.0);
    dqdl_.assign(vf_.size() * 3, 0.0);
    if (real_) {
        rg_.D = g_->dim;
        rg_.M = M_;
        for (int a = 0; a < 3; a++) rg_.tn[a] = g_->tnx[a];
        rg_.st[2] = 1;
        rg_.st[1] = rg_.tn[2];
        rg_.st[0] = (long long)rg_.tn[1] * rg_.tn[2];
        int off = 0;
        for (int a = 0; a < rg_.D; a++) {
            rg_.ax[a].n = g_->nx[a];
            rg_.ax[a].iax = a + (3 - rg_.D);
            rg_.ax[a].periodic = bc_[2 * a] == BC_PERIODIC;
            rg_.ax[a].off = off;
            off += RT_ROWS * g_->nx[a] + 2;
        }
        rg_.tab_stride = off;
        rg_.half_ds = 0.5 * ds_;
        rtab_.allocate(sizeof(double) * (size_t)off * n_mon_);
        build_real_tables();
    } else {
        bt_.build(*g_, bsq_, ds_);
    }

    // partition segment: first block of each polymer in block order (discrete: with
    // n_left != 0), CpuComputationContinuous.cpp:67-91, CpuComputationDiscrete.cpp:85-115
    part_block_.assign(vf_.size(), -1);
    int cur = 0;
    for (int b = 0; b < (int)blocks_.size() && cur < (int)vf_.size(); b++) {
        if (blocks_[b].polymer != cur) continue;
        if (model_ == FTS_DISCRETE && blocks_[b].nl == 0) continue;
        part_block_[cur++] = b;
    }

    for (auto& jobs : slices_) build_slice(jobs);
    build_stress_terms();
}

template <typename T>
void Solver<T>::emit_slab(int mode, const std::vector<SlabItem<T>>& items) {
    for (size_t i = 0; i < items.size(); i += MAX_BATCH) {
        SlabBatch<T> b{};
        b.n = (int)std::min<size_t>(MAX_BATCH, items.size() - i);
        for (int k = 0; k < b.n; k++) b.it[k] = items[i + k];
        launches_.push_back({mode == SLAB_FWD ? L_SLAB_FWD : mode == SLAB_INV1 ? L_SLAB_INV1 : L_SLAB_INV2,
                             (int)slabs_.size()});
        slabs_.push_back(b);
    }
}

template <typename T>
void Solver<T>::emit_x(const std::vector<XItem<T>>& items) {
    for (size_t i = 0; i < items.size(); i += MAX_BATCH) {
        XBatch<T> b{};
        b.n = (int)std::min<size_t>(MAX_BATCH, items.size() - i);
        for (int k = 0; k < b.n; k++) b.it[k] = items[i + k];
        launches_.push_back({L_XPASS, (int)xs_.size()});
        xs_.push_back(b);
    }
}

template <typename T>
void Solver<T>::emit_init(const std::vector<InitItem<T>>& items) {
    for (size_t i = 0; i < items.size(); i += INIT_BATCH) {
        InitBatch<T> b{};
        b.n = (int)std::min<size_t>(INIT_BATCH, items.size() - i);
        for (int k = 0; k < b.n; k++) b.it[k] = items[i + k];
        launches_.push_back({L_INIT, (int)inits_.size()});
        inits_.push_back(b);
    }
}

template <typename T>
void Solver<T>::emit_real(const std::vector<RealItem>& items) {
    for (size_t i = 0; i < items.size(); i += MAX_BATCH) {
        RealBatch b{};
        b.n = (int)std::min<size_t>(MAX_BATCH, items.size() - i);
        for (int k = 0; k < b.n; k++) b.it[k] = items[i + k];
        launches_.push_back({L_REAL, (int)reals_.size()});
        reals_.push_back(b);
    }
}

// Per monomer and axis: FiniteDifference.cpp:6-101 rows (l, d, h), then the
// factors of the system that is solved along the axis. Periodic axes: the cyclic
// system A = T + u v^T with gamma = -d[0], T = A with d[0] - gamma and
// d[n-1] - h[n-1] l[0] / gamma, u = (gamma, 0, .., h[n-1]), v = (1, 0, .., l[0] / gamma);
// z = T^-1 u and 1 / (1 + v.z) are tabulated.
template <typename T>
void Solver<T>::build_real_tables() {
    std::vector<double> tab((size_t)rg_.tab_stride * n_mon_, 0.0);
    for (int m = 0; m < n_mon_; m++) {
        for (int a = 0; a < rg_.D; a++) {
            const RealAxis& A = rg_.ax[a];
            const int n = A.n;
            double* t = tab.data() + (size_t)m * rg_.tab_stride + A.off;
            double *l = t + RT_L * n, *d = t + RT_D * n, *h = t + RT_H * n;
            double *inv = t + RT_INV * n, *c = t + RT_C * n, *z = t + RT_Z * n;
            const double f = bsq_[m] * ds_ / (g_->dx[a] * g_->dx[a] * 6.0);
            for (int i = 0; i < n; i++) {
                l[i] = -0.5 * f;
                d[i] = 1.0 + f;
                h[i] = -0.5 * f;
            }
            const int lo = bc_[2 * a], hi = bc_[2 * a + 1];
            if (lo == BC_REFLECTING) {
                d[0] += l[0];
                l[0] = 0.0;
            } else if (lo == BC_ABSORBING) {
                l[0] = 0.0;
            }
            if (hi == BC_REFLECTING) {
                d[n - 1] += h[n - 1];
                h[n - 1] = 0.0;
            } else if (hi == BC_ABSORBING) {
                h[n - 1] = 0.0;
            }
            std::vector<double> dt(d, d + n), u(n, 0.0);
            double gamma = 0.0;
            if (A.periodic) {
                gamma = -d[0];
                dt[0] -= gamma;
                dt[n - 1] -= h[n - 1] * l[0] / gamma;
                u[0] += gamma;
                u[n - 1] += h[n - 1];
            }
            inv[0] = 1.0 / dt[0];
            for (int i = 1; i < n; i++) {
                c[i - 1] = h[i - 1] * inv[i - 1];
                inv[i] = 1.0 / (dt[i] - l[i] * c[i - 1]);
            }
            if (A.periodic) {
                std::vector<double> y(n);
                y[0] = u[0] * inv[0];
                for (int i = 1; i < n; i++) y[i] = (u[i] - l[i] * y[i - 1]) * inv[i];
                z[n - 1] = y[n - 1];
                for (int i = n - 2; i >= 0; i--) z[i] = y[i] - c[i] * z[i + 1];
                const double v = l[0] / gamma;
                t[RT_ROWS * n] = v;
                t[RT_ROWS * n + 1] = 1.0 / (1.0 + z[0] + v * z[n - 1]);
            }
        }
    }
    FTS_HIP_CHECK(hipMemcpy(rtab_.get(), tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
}

// Bytes one step must move per item, each launch reading and writing its operands
// once: pre, on a run's first step (q_in, w -> q_e), first sweep (q_e -> acc),
// further sweeps (acc, q_e -> acc), last sweep (acc, q_e, w, mask -> q_out, and
// q_e of the next step unless this is the run's last).
static double real_step_bytes(const RealGeom& g, const RealItem& it) {
    const double e = sizeof(double) * (double)g.M;
    const double masked = it.mask ? 1.0 : 0.0;
    const double pre = it.need_pre ? 3 * e : 0.0;
    const double out = (it.write_qe ? 1.0 : 0.0) * e;
    if (g.D == 1) return pre + (3 + masked) * e + out;
    return pre + 2 * e + 3 * e * (g.D - 2) + (4 + masked) * e + out;
}

template <typename T>
void Solver<T>::launch_real(const RealBatch& b) {
    KernelTimer& timer = g_->ctx->timer;
    hipEvent_t start = timer.begin(g_->stream(), FTS_KERNEL_REAL);
    bool any_pre = false;
    for (int k = 0; k < b.n; k++) any_pre = any_pre || b.it[k].need_pre;
    if (any_pre) {
        const int pb = std::min(1024, ceil_div(M_, 256));
        hipLaunchKernelGGL(real_pre_kernel, dim3(pb, b.n), dim3(256), 0, g_->stream(), b, rg_);
        FTS_HIP_CHECK(hipGetLastError());
    }
    const double* tab = rtab_.as<double>();
    for (int a = 0; a < rg_.D; a++) {
        const RealAxis& A = rg_.ax[a];
        const bool first = a == 0, last = a == rg_.D - 1;
        if (A.n <= 64 * RT_LMAX && !real_plain_) {
            RealTile t{};
            t.P = 1;
            while (t.P * RT_LMAX < A.n) t.P <<= 1;
            t.Lc = ceil_div(A.n, t.P);
            t.R = RT_THREADS / t.P;
            const bool inner = A.iax == 2;
            int blocks;
            if (inner) {
                blocks = ceil_div(M_ / A.n, t.R);
            } else {
                t.ktiles = ceil_div(rg_.tn[2], t.R);
                blocks = rg_.tn[A.iax == 0 ? 1 : 0] * t.ktiles;
            }
            const dim3 grid(blocks, b.n);
            const size_t lds = sizeof(double) * (size_t)(t.R + 4) * A.n;
#define FTS_REAL_TILED(F, L, I)                                                                                \
    hipLaunchKernelGGL((real_tiled_kernel<F, L, I>), grid, dim3(RT_THREADS), lds, g_->stream(), b, rg_, tab, a, t)
            if (inner) {
                if (first && last) FTS_REAL_TILED(true, true, true);
                else if (first) FTS_REAL_TILED(true, false, true);
                else if (last) FTS_REAL_TILED(false, true, true);
                else FTS_REAL_TILED(false, false, true);
            } else {
                if (first && last) FTS_REAL_TILED(true, true, false);
                else if (first) FTS_REAL_TILED(true, false, false);
                else if (last) FTS_REAL_TILED(false, true, false);
                else FTS_REAL_TILED(false, false, false);
            }
#undef FTS_REAL_TILED
            FTS_HIP_CHECK(hipGetLastError());
            continue;
        }
        const int lines = M_ / rg_.ax[a].n;
        const dim3 grid(ceil_div(lines, 256), b.n);
        if (first && last) hipLaunchKernelGGL((real_sweep_kernel<true, true>), grid, dim3(256), 0, g_->stream(), b, rg_, tab, a);
        else if (first) hipLaunchKernelGGL((real_sweep_kernel<true, false>), grid, dim3(256), 0, g_->stream(), b, rg_, tab, a);
        else if (last) hipLaunchKernelGGL((real_sweep_kernel<false, true>), grid, dim3(256), 0, g_->stream(), b, rg_, tab, a);
        else hipLaunchKernelGGL((real_sweep_kernel<false, false>), grid, dim3(256), 0, g_->stream(), b, rg_, tab, a);
        FTS_HIP_CHECK(hipGetLastError());
    }
    if (start) {
        double bytes = 0.0;
        for (int k = 0; k < b.n; k++) bytes += real_step_bytes(rg_, b.it[k]);
        timer.end(g_->stream(), start, FTS_KERNEL_REAL, bytes);
    }
}

template <typename T>
void Solver<T>::build_slice(const std::vector<std::array<int, 3>>& jobs) {
    const T ds = (T)ds_;
    // ---- initial slices of propagators starting in this slice
    // rounds of at most INIT_MAX_SRC dependencies; later rounds accumulate
    std::vector<std::vector<InitItem<T>>> rounds;
    auto add_init = [&](int round, const InitItem<T>& it) {
        if ((int)rounds.size() <= round) rounds.resize(round + 1);
        rounds[round].push_back(it);
    };
    struct FftInit {
        int p, slot;
        T* src;
        int table;
    };
    std::vector<FftInit> fft_inits;
    for (size_t s = 0; s < jobs.size(); s++) {
        const int p = jobs[s][0], from = jobs[s][1];
        if (from != 0) continue;
        const Prop& pr = props_[p];
        const int m = pr.monomer;
        if (model_ == FTS_CONTINUOUS) {
            InitItem<T> it{};
            it.dst = q(p, 0);
            if (pr.kind == FTS_LEAF || pr.kind == FTS_LEAF_QINIT) {
                it.op = pr.kind == FTS_LEAF ? INIT_ONE : INIT_QINIT;
                if (pr.kind == FTS_LEAF_QINIT) it.qinit = qinit_.as<double>() + (size_t)pr.qinit * M_;
                it.mask = mask();
                add_init(0, it);
            } else {
                const int op = pr.kind == FTS_AGGREGATED ? INIT_SUM : INIT_PROD;
                const int nd = (int)pr.deps.size();
                for (int r = 0; r * INIT_MAX_SRC < nd; r++) {
                    InitItem<T> c{};
                    c.dst = q(p, 0);
                    c.op = op;
                    c.accumulate = r > 0;
                    c.n_src = std::min(INIT_MAX_SRC, nd - r * INIT_MAX_SRC);
                    for (int k = 0; k < c.n_src; k++) {
                        const auto& d = pr.deps[r * INIT_MAX_SRC + k];
                        c.src[k] = q(d[0], d[1]);
                        c.coef[k] = (T)d[2];
                    }
                    if ((r + 1) * INIT_MAX_SRC >= nd) c.mask = mask();
                    add_init(r, c);
                }
            }
        } else {
            if (pr.kind == FTS_LEAF || pr.kind == FTS_LEAF_QINIT) {
                if (pr.max_n < 1) continue;
                InitItem<T> it{};
                it.dst = q(p, 1);
                it.op = pr.kind == FTS_LEAF ? INIT_ONE : INIT_QINIT;
                if (pr.kind == FTS_LEAF_QINIT) it.qinit = qinit_.as<double>() + (size_t)pr.qinit * M_;
                it.w = ew(m);
                it.wcoef = ew_coef();
                it.mask = mask();
                add_init(0, it);
            } else {
                const bool agg = pr.kind == FTS_AGGREGATED;
                const bool via_half = !agg || pr.deps[0][1] == 0;
                T* dst = via_half ? h(p, 0) : q(p, 1);
                const int nd = (int)pr.deps.size();
                for (int r = 0; r * INIT_MAX_SRC < nd; r++) {
                    InitItem<T> c{};
                    c.dst = dst;
                    c.op = agg ? INIT_SUM : INIT_PROD;
                    c.accumulate = r > 0;
                    c.n_src = std::min(INIT_MAX_SRC, nd - r * INIT_MAX_SRC);
                    for (int k = 0; k < c.n_src; k++) {
                        const auto& d = pr.deps[r * INIT_MAX_SRC + k];
                        // junction: half-bond slices of the deps; aggregated: q(n) or junction start
                        c.src[k] = (!agg || d[1] == 0) ? h(d[0], d[1]) : q(d[0], d[1]);
                        c.coef[k] = (T)d[2];
                    }
                    add_init(r, c);
                }
                if (pr.max_n > 0) fft_inits.push_back({p, (int)s, dst, via_half ? half(m) : full(m)});
            }
        }
    }
    for (auto& r : rounds) emit_init(r);
    if (!fft_inits.empty()) {
        // q[1] = E * mask * IFFT[B FFT[src]] with B the half bond (junction) or full bond
        std::vector<SlabItem<T>> fw, inv;
        std::vector<XItem<T>> xi;
        for (auto& f : fft_inits) {
            SlabItem<T> a{};
            a.qin = f.src;
            a.c0 = slot(f.slot);
            a.n_out = 1;
            fw.push_back(a);
            xi.push_back({slot(f.slot), f.table});
            SlabItem<T> b{};
            b.c0 = slot(f.slot);
            b.w = ew(props_[f.p].monomer);
            b.post0 = ew_coef();
            b.mask = mask();
            b.qout = q(f.p, 1);
            inv.push_back(b);
        }
        emit_slab(SLAB_FWD, fw);
        emit_x(xi);
        emit_slab(SLAB_INV1, inv);
    }

    // ---- steps: jobs advance in lockstep, aligned at the end of the slice
    struct Run {
        int p, slot, n0, cnt;
    };
    std::vector<Run> runs;
    int D = 0;
    for (size_t s = 0; s < jobs.size(); s++) {
        const int p = jobs[s][0], from = jobs[s][1], to = jobs[s][2];
        if (props_[p].max_n == 0) continue;
        const int n0 = model_ == FTS_CONTINUOUS ? from : std::max(1, from);
        const int cnt = to - n0;
        if (cnt <= 0) continue;
        runs.push_back({p, (int)s, n0, cnt});
        D = std::max(D, cnt);
    }
    for (int t = 0; t < D; t++) {
        std::vector<SlabItem<T>> fw, inv1, inv2;
        std::vector<XItem<T>> x1, x2;
        std::vector<RealItem> ri;
        for (auto& r : runs) {
            const int start = D - r.cnt;
            if (t < start) continue;
            const int n = r.n0 + (t - start);
            const int m = props_[r.p].monomer;
            const bool last = t == D - 1;
            if (real_) {
                if constexpr (std::is_same<T, double>::value) {
                    RealItem it{};
                    it.qin = q(r.p, n);
                    it.qe = reinterpret_cast<double*>(slot(2 * r.slot));
                    it.acc = reinterpret_cast<double*>(slot(2 * r.slot + 1));
                    it.qout = q(r.p, n + 1);
                    it.w = w(m);
                    it.mask = mask();
                    it.monomer = m;
                    it.need_pre = t == start;
                    it.write_qe = !last;
                    ri.push_back(it);
                }
            } else if (model_ == FTS_CONTINUOUS) {
                T2* c0 = slot(2 * r.slot);
                T2* c1 = slot(2 * r.slot + 1);
                if (t == start) {
                    SlabItem<T> a{};
                    a.qin = q(r.p, n);
                    a.c0 = c0;
                    a.c1 = c1;
                    a.w = w(m);
                    a.pre0 = ds / 2;
                    a.pre1 = ds / 4;
                    a.n_out = 2;
                    fw.push_back(a);
                }
                x1.push_back({c0, full(m)});
                x1.push_back({c1, half(m)});
                SlabItem<T> b{};
                b.c0 = c1;
                b.w = w(m);
                b.post0 = ds / 2;
                b.forward = 1;
                inv1.push_back(b);
                x2.push_back({c1, half(m)});
                SlabItem<T> c{};
                c.c0 = c0;
                c.c1 = c1;
                c.w = w(m);
                c.post0 = ds / 2;
                c.post1 = ds / 4;
                c.mask = mask();
                c.qout = q(r.p, n + 1);
                c.forward = last ? 0 : 1;
                c.pre0 = ds / 2;
                c.pre1 = ds / 4;
                inv2.push_back(c);
            } else {
                T2* c0 = slot(r.slot);
                if (t == start) {
                    SlabItem<T> a{};
                    a.qin = q(r.p, n);
                    a.c0 = c0;
                    a.n_out = 1;
                    fw.push_back(a);
                }
                x1.push_back({c0, full(m)});
                SlabItem<T> b{};
                b.c0 = c0;
                b.w = ew(m);
                b.post0 = ew_coef();
                b.mask = mask();
                b.qout = q(r.p, n + 1);
                b.forward = last ? 0 : 1;
                inv1.push_back(b);
            }
        }
        if (real_) {
            emit_real(ri);
            continue;
        }
        emit_slab(SLAB_FWD, fw);
        emit_x(x1);
        emit_slab(SLAB_INV1, inv1);
        if (model_ == FTS_CONTINUOUS) {
            emit_x(x2);
            emit_slab(SLAB_INV2, inv2);
        }
    }

    // ---- discrete half-bond slices at junction ends reached in this slice
    if (model_ == FTS_DISCRETE) {
        std::vector<std::pair<int, int>> hs;
        for (auto& jb : jobs) {
            const int p = jb[0], from = jb[1], to = jb[2];
            for (int n : props_[p].je)
                if (n > from && n <= to && n >= 1 && props_[p].max_n > 0) hs.push_back({p, n});
        }
        for (size_t i = 0; i < hs.size(); i += n_ws_) {
            std::vector<SlabItem<T>> fw, inv;
            std::vector<XItem<T>> xi;
            for (size_t k = i; k < std::min(hs.size(), i + n_ws_); k++) {
                const int p = hs[k].first, n = hs[k].second, sl = (int)(k - i);
                SlabItem<T> a{};
                a.qin = q(p, n);
                a.c0 = slot(sl);
                a.n_out = 1;
                fw.push_back(a);
                xi.push_back({slot(sl), half(props_[p].monomer)});
                SlabItem<T> b{};
                b.c0 = slot(sl);
                b.qout = h(p, n);
                inv.push_back(b);
            }
            emit_slab(SLAB_FWD, fw);
            emit_x(xi);
            emit_slab(SLAB_INV1, inv);
        }
    }
}

template <typename T>
void Solver<T>::build_stress_terms() {
    // CpuComputationContinuous.cpp:671-693 and CpuComputationDiscrete.cpp:905-973, in block order
    for (int b = 0; b < (int)blocks_.size(); b++) {
        const Block& bk = blocks_[b];
        if (bk.nr == 0) continue;
        const double a2 = bsq_[bk.monomer];
        for (int n = 0; n <= bk.nr; n++) {
            STerm t{};
            t.block = b;
            if (model_ == FTS_CONTINUOUS) {
                const int N = bk.nr;
                double s;
                if (N == 1) s = 0.5;
                else if (N == 3) s = (n == 0 || n == 3) ? 3.0 / 8.0 : 9.0 / 8.0;
                else if (N % 2 == 0) s = (n == 0 || n == N) ? 1.0 / 3.0 : (n % 2 == 1 ? 4.0 / 3.0 : 2.0 / 3.0);
                else if (n == 0) s = 1.0 / 3.0;
                else if (n <= N - 4) s = n % 2 == 1 ? 4.0 / 3.0 : 2.0 / 3.0;
                else if (n == N - 3) s = 3.0 / 8.0 + 1.0 / 3.0;
                else if (n == N - 2 || n == N - 1) s = 9.0 / 8.0;
                else s = 3.0 / 8.0;
                t.x1p = bk.left;
                t.x1n = bk.nl - n;
                t.x2p = bk.right;
                t.x2n = n;
                t.bsq = a2;
                t.table = -1;
                t.weight = s * bk.rep;
            } else {
                if (n == bk.nl) {
                    if (props_[bk.left].deps.empty()) continue;
                    t.x1p = bk.left; t.x1n = 0; t.x1h = 1;
                    t.x2p = bk.right; t.x2n = bk.nr;
                    t.bsq = 0.5 * a2;
                    t.table = half(bk.monomer);
                } else if (n == 0 && bk.nl == bk.nr) {
                    if (props_[bk.right].deps.empty()) continue;
                    t.x1p = bk.left; t.x1n = bk.nl;
                    t.x2p = bk.right; t.x2n = 0; t.x2h = 1;
                    t.bsq = 0.5 * a2;
                    t.table = half(bk.monomer);
                } else if (n == 0) {
                    continue;
                } else {
                    t.x1p = bk.left; t.x1n = bk.nl - n;
                    t.x2p = bk.right; t.x2n = n;
                    t.bsq = a2;
                    t.table = full(bk.monomer);
                }
                t.weight = bk.rep;
            }
            sterms_.push_back(t);
        }
    }
    stress_out_.allocate(sizeof(double) * 3 * std::max<size_t>(1, sterms_.size()));
    if (g_->wave_local && !real_ && !sterms_.empty())
        stress_part_.allocate(sizeof(double) * 3 * (size_t)std::min(n_ws_ / 2, MAX_BATCH) *
                              g_->stress_pencil_blocks_max());
}

template <typename T>
void Solver<T>::run_launches() {
    const T* bx = bt_.x<T>();
    const T* by = bt_.y<T>();
    const T* bz = bt_.z<T>();
    const int ib = std::min(1024, ceil_div(M_, 256));
    for (const Launch& L : launches_) {
        switch (L.type) {
            case L_INIT: {
                const InitBatch<T>& b = inits_[L.idx];
                hipLaunchKernelGGL((init_kernel<T>), dim3(ib, b.n), dim3(256), 0, g_->stream(), b, M_);
                FTS_HIP_CHECK(hipGetLastError());
                break;
            }
            case L_SLAB_FWD: g_->launch_slab<T>(SLAB_FWD, slabs_[L.idx]); break;
            case L_SLAB_INV1: g_->launch_slab<T>(SLAB_INV1, slabs_[L.idx]); break;
            case L_SLAB_INV2: g_->launch_slab<T>(SLAB_INV2, slabs_[L.idx]); break;
            case L_XPASS: g_->launch_xpass<T>(xs_[L.idx], bx, by, bz); break;
            case L_REAL: launch_real(reals_[L.idx]); break;
        }
    }
}

template <typename T>
typename Solver<T>::GraphKey Solver<T>::graph_key(const double* const* w, const double* const* q_init) const {
    GraphKey k;
    for (int m = 0; m < n_mon_; m++) k.ptrs.push_back(w ? w[m] : nullptr);
    for (int i = 0; i < n_qinit_; i++) k.ptrs.push_back(q_init ? q_init[i] : nullptr);
    k.volume = g_->volume;
    const int p = std::is_same<T, float>::value ? 1 : 0;
    k.bd = g_->slab_bd[p];
    k.bd_fwd = g_->slab_bd_mode[p][SLAB_FWD];
    k.bd_inv2 = g_->slab_bd_mode[p][SLAB_INV2];
    k.nl = g_->pencil_nl[p];
    return k;
}

// sampled launch events stay out of graphs: with kernel timing on, calls run eagerly
template <typename T>
bool Solver<T>::graph_ok() const {
    return graphs_ && !g_->ctx->timer.active();
}

template <typename T>
void Solver<T>::drop_graphs() {
    for (hipGraphExec_t* e : {&gprop_, &gconc_}) {
        if (*e) (void)hipGraphExecDestroy(*e);
        *e = nullptr;
    }
    kprop_ = GraphKey{};
    kconc_ = GraphKey{};
}

template <typename T>
template <typename F>
void Solver<T>::replay_or_capture(hipGraphExec_t& exec, GraphKey& have, const GraphKey& want, F&& issue) {
    hipStream_t st = g_->stream();
    if (!exec || !(have == want)) {
        if (exec) FTS_HIP_CHECK(hipGraphExecDestroy(exec));
        exec = nullptr;
        FTS_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipGraph_t graph = nullptr;
        try {
            issue();
        } catch (...) {
            (void)hipStreamEndCapture(st, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        FTS_HIP_CHECK(hipStreamEndCapture(st, &graph));
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        FTS_HIP_CHECK(e);
        have = want;
    }
    FTS_HIP_CHECK(hipGraphLaunch(exec, st));
}

template <typename T>
void Solver<T>::compute_propagators(const double* const* wf, const double* const* q_init, bool on_device) {
    for (int m = 0; m < n_mon_; m++) FTS_REQUIRE(wf[m] != nullptr, "missing field for a monomer type");
    for (int i = 0; i < n_qinit_; i++) FTS_REQUIRE(q_init && q_init[i], "missing q_init field");
    if (on_device && graph_ok())
        replay_or_capture(gprop_, kprop_, graph_key(wf, q_init), [&] { issue_propagators(wf, q_init, true); });
    else
        issue_propagators(wf, q_init, on_device);
    Q_valid_ = false;
}

template <typename T>
void Solver<T>::issue_propagators(const double* const* wf, const double* const* q_init, bool on_device) {
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const int cb = std::min(1024, ceil_div(M_, 256));
    for (int m = 0; m < n_mon_; m++) {
        FTS_REQUIRE(wf[m] != nullptr, "missing field for a monomer type");
        if (std::is_same<T, double>::value) {
            FTS_HIP_CHECK(hipMemcpyAsync(w64_.as<double>() + (size_t)m * M_, wf[m], sizeof(double) * M_, kind,
                                         g_->stream()));
        } else if (on_device) {
            continue;   // converted below, straight from the caller's device fields
        } else {
            FTS_HIP_CHECK(hipMemcpyAsync(w64_.as<double>() + (size_t)m * M_, wf[m], sizeof(double) * M_, kind,
                                         g_->stream()));
            hipLaunchKernelGGL((convert_kernel<T>), dim3(cb), dim3(256), 0, g_->stream(),
                               w64_.as<double>() + (size_t)m * M_, wT_ + (size_t)m * M_, M_);
        }
    }
    if (!std::is_same<T, double>::value && on_device) {
        for (int m0 = 0; m0 < n_mon_; m0 += MAX_BATCH) {
            ConvertBatch cv{};
            cv.n = std::min(MAX_BATCH, n_mon_ - m0);
            for (int k = 0; k < cv.n; k++) cv.src[k] = wf[m0 + k];
            hipLaunchKernelGGL((convert_batch_kernel<T>), dim3(cb, cv.n), dim3(256), 0, g_->stream(), cv,
                               wT_ + (size_t)m0 * M_, M_);
        }
    }
    if (!std::is_same<T, double>::value) FTS_HIP_CHECK(hipGetLastError());
    if (table()) {
        hipLaunchKernelGGL((exp_table_kernel<T>), dim3(std::min(1024, ceil_div(n_mon_ * M_, 256))), dim3(256), 0,
                           g_->stream(), wT_, (T)ds_, ew_.as<T>(), n_mon_ * M_);
        FTS_HIP_CHECK(hipGetLastError());
    }
    for (int i = 0; i < n_qinit_; i++) {
        FTS_REQUIRE(q_init && q_init[i], "missing q_init field");
        FTS_HIP_CHECK(hipMemcpyAsync(qinit_.as<double>() + (size_t)i * M_, q_init[i], sizeof(double) * M_, kind,
                                     g_->stream()));
    }
    run_launches();

    // Q_p = <q_L[n_L], q_R[0 or 1]>/n_agg/V  (discrete: /E)
    DotBatch<T> db{};
    std::vector<int> target;
    auto flush = [&]() {
        if (!db.n) return;
        bool consecutive = true;
        for (int k = 1; k < db.n; k++) consecutive = consecutive && target[k] == target[0] + k;
        if (consecutive) {
            g_->launch_dots<T>(db, Qd_.as<double>() + target[0]);
        } else {
            g_->launch_dots<T>(db, g_->scalars.as<double>());
            for (int k = 0; k < db.n; k++)
                FTS_HIP_CHECK(hipMemcpyAsync(Qd_.as<double>() + target[k], g_->scalars.as<double>() + k,
                                             sizeof(double), hipMemcpyDeviceToDevice, g_->stream()));
        }
        db.n = 0;
        target.clear();
    };
    for (size_t p = 0; p < vf_.size(); p++) {
        const int b = part_block_[p];
        FTS_REQUIRE(b >= 0, "polymer without a partition segment");
        const Block& bk = blocks_[b];
        const int nagg = bk.nvu / bk.rep;
        DotItem<T> it{};
        it.a = q(bk.left, bk.nl);
        it.b = q(bk.right, model_ == FTS_CONTINUOUS ? 0 : 1);
        if (model_ == FTS_DISCRETE) {
            it.w = w(bk.monomer);
            it.wcoef = (T)ds_;
        }
        it.scale = 1.0 / nagg / g_->volume;
        db.it[db.n++] = it;
        target.push_back((int)p);
        if (db.n == MAX_BATCH) flush();
    }
    flush();
    Q_valid_ = false;
}

template <typename T>
void Solver<T>::compute_concentrations() {
    if (graph_ok() && gprop_)
        replay_or_capture(gconc_, kconc_, graph_key(nullptr, nullptr), [&] { issue_concentrations(); });
    else
        issue_concentrations();
    Q_valid_ = false;
}

template <typename T>
void Solver<T>::issue_concentrations() {
    const int gb = std::min(1024, ceil_div(M_, 256));
    for (size_t b0 = 0; b0 < blocks_.size(); b0 += MAX_BATCH) {
        PhiBatch<T> pb{};
        pb.n = (int)std::min<size_t>(MAX_BATCH, blocks_.size() - b0);
        for (int k = 0; k < pb.n; k++) {
            const Block& bk = blocks_[b0 + k];
            PhiItem<T>& it = pb.it[k];
            it.qL = q(bk.left, 0);
            it.qR = q(bk.right, 0);
            it.NL = bk.nl;
            it.NR = bk.nr;
            it.polymer = bk.polymer;
            it.norm = ds_ * vf_[bk.polymer] / alpha_[bk.polymer] * bk.rep;
            it.w = w(bk.monomer);
            it.wcoef = (T)ds_;
            it.out = phib_.as<double>() + (b0 + k) * (size_t)M_;
        }
        if (model_ == FTS_CONTINUOUS)
            hipLaunchKernelGGL((phi_kernel<T, false>), dim3(gb, pb.n), dim3(256), 0, g_->stream(), pb, M_,
                               Qd_.as<double>());
        else
            hipLaunchKernelGGL((phi_kernel<T, true>), dim3(gb, pb.n), dim3(256), 0, g_->stream(), pb, M_,
                               Qd_.as<double>());
        FTS_HIP_CHECK(hipGetLastError());
    }
    for (size_t s = 0; s < svf_.size(); s++) {
        DotBatch<T> db{};
        db.n = 1;
        db.it[0] = DotItem<T>{nullptr, nullptr, w(smon_[s]), (T)-ds_, 1.0 / g_->volume};
        const int slot_q = (int)(vf_.size() + s);
        g_->launch_dots<T>(db, Qd_.as<double>() + slot_q);
        hipLaunchKernelGGL((solvent_phi_kernel<T>), dim3(gb), dim3(256), 0, g_->stream(), w(smon_[s]), ds_, svf_[s],
                           Qd_.as<double>() + slot_q, M_, phis_.as<double>() + s * (size_t)M_);
        FTS_HIP_CHECK(hipGetLastError());
    }
    Q_valid_ = false;
}

template <typename T>
void Solver<T>::compute_stress() {
    // CpuSolverReal.cpp:435-450
    if (real_) throw Error("Currently, the real-space method does not support stress computation.");
    const int K = std::min(n_ws_ / 2, MAX_BATCH);   // terms per chunk: 2K spectra, K reduction items
    std::vector<double> seg(sterms_.size() * 3, 0.0);
    StressGeom sg{};
    sg.n0 = g_->tnx[0];
    sg.n1 = g_->tnx[1];
    sg.nz = g_->tnx[2];
    sg.nzc = g_->nzc;
    sg.sp = g_->spec_pitch<T>();
    const double PI = 3.14159265358979323846;
    sg.f0 = std::pow(2 * PI / (g_->tnx[0] * g_->tdx[0]), 2);
    sg.f1 = std::pow(2 * PI / (g_->tnx[1] * g_->tdx[1]), 2);
    sg.f2 = std::pow(2 * PI / (g_->tnx[2] * g_->tdx[2]), 2);
    for (size_t t0 = 0; t0 < sterms_.size(); t0 += K) {
        const int nt = (int)std::min<size_t>(K, sterms_.size() - t0);
        std::vector<SlabItem<T>> fw;
        std::vector<XItem<T>> xi;
        StressBatch<T> sb{};
        sb.n = nt;
        for (int k = 0; k < nt; k++) {
            const STerm& t = sterms_[t0 + k];
            SlabItem<T> a{};
            a.qin = slice_ptr(t.x1p, t.x1n, t.x1h);
            a.c0 = slot(2 * k);
            a.n_out = 1;
            fw.push_back(a);
            a.qin = slice_ptr(t.x2p, t.x2n, t.x2h);
            a.c0 = slot(2 * k + 1);
            fw.push_back(a);
            xi.push_back({slot(2 * k), -1});
            xi.push_back({slot(2 * k + 1), -1});
            sb.it[k] = StressItem<T>{slot(2 * k), slot(2 * k + 1), t.table, t.bsq};
        }
        const bool fused = stress_part_.get() != nullptr;
        for (size_t i = 0; i < fw.size(); i += MAX_BATCH) {
            SlabBatch<T> b{};
            b.n = (int)std::min<size_t>(MAX_BATCH, fw.size() - i);
            for (int k = 0; k < b.n; k++) b.it[k] = fw[i + k];
            g_->launch_slab<T>(SLAB_FWD, b);
            if (fused) continue;
            XBatch<T> x{};
            x.n = b.n;
            for (int k = 0; k < b.n; k++) x.it[k] = xi[i + k];
            g_->launch_xpass<T>(x, nullptr, nullptr, nullptr);
        }
        if (fused) {
            // wave-local grids: x transforms and reduction in one pass, no spectrum written
            const int nb = g_->launch_stress_pencil<T>(sb, sg, bt_.x<T>(), bt_.y<T>(), bt_.z<T>(),
                                                       stress_part_.as<double>());
            hipLaunchKernelGGL(stress_stage2_n, dim3(nt), dim3(256), 0, g_->stream(), stress_part_.as<double>(), nb,
                               stress_out_.as<double>() + t0 * 3);
            FTS_HIP_CHECK(hipGetLastError());
            continue;
        }
        hipLaunchKernelGGL((stress_stage1<T>), dim3(RED_BLOCKS, nt), dim3(RED_THREADS), 0, g_->stream(), sb, sg,
                           bt_.x<T>(), bt_.y<T>(), bt_.z<T>(), g_->partials.as<double>());
        FTS_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(stress_stage2, dim3(nt), dim3(RED_BLOCKS), 0, g_->stream(), g_->partials.as<double>(),
                           stress_out_.as<double>() + t0 * 3);
        FTS_HIP_CHECK(hipGetLastError());
    }
    if (!sterms_.empty())
        FTS_HIP_CHECK(hipMemcpyAsync(seg.data(), stress_out_.get(), sizeof(double) * seg.size(),
                                     hipMemcpyDeviceToHost, g_->stream()));
    FTS_HIP_CHECK(hipStreamSynchronize(g_->stream()));
    // accumulate in the reference order: per block, then per polymer, then normalise
    const int dim = g_->dim;
    std::vector<double> bdq(blocks_.size() * 3, 0.0);
    for (size_t i = 0; i < sterms_.size(); i++)
        for (int d = 0; d < dim; d++) bdq[sterms_[i].block * 3 + d] += seg[i * 3 + (d + 3 - dim)] * sterms_[i].weight;
    std::fill(dqdl_.begin(), dqdl_.end(), 0.0);
    for (size_t b = 0; b < blocks_.size(); b++)
        for (int d = 0; d < dim; d++) dqdl_[blocks_[b].polymer * 3 + d] += bdq[b * 3 + d];
    for (size_t p = 0; p < vf_.size(); p++)
        for (int d = 0; d < dim; d++) dqdl_[p * 3 + d] /= -3.0 * g_->lx[d] * M_ * M_ / ds_;
}

template <typename T>
double Solver<T>::partition(int p) {
    FTS_REQUIRE(p >= 0 && p < (int)vf_.size(), "polymer index out of range");
    if (!Q_valid_) {
        FTS_HIP_CHECK(hipMemcpyAsync(Qh_.data(), Qd_.get(), sizeof(double) * Qh_.size(), hipMemcpyDeviceToHost,
                                     g_->stream()));
        FTS_HIP_CHECK(hipStreamSynchronize(g_->stream()));
        Q_valid_ = true;
    }
    return Qh_[p];
}

template <typename T>
double Solver<T>::solvent_partition(int s) {
    FTS_REQUIRE(s >= 0 && s < (int)svf_.size(), "solvent index out of range");
    partition(0);
    return Qh_[vf_.size() + s];
}

template <typename T>
void Solver<T>::dq_dl(int p, double* out) {
    FTS_REQUIRE(p >= 0 && p < (int)vf_.size(), "polymer index out of range");
    for (int d = 0; d < g_->dim; d++) out[d] = dqdl_[p * 3 + d];
}

template <typename T>
void Solver<T>::sum_concentrations_replicas(int n, const int* kind, const int* idx, const double* scale, double* out,
                                            int n_rep, int block_step, int solvent_step, size_t out_step) {
    FTS_REQUIRE(n_rep >= 1 && n_rep <= 65535, "replica count out of range");
    const int gb = std::min(1024, ceil_div(M_, 256));
    SumBatch sb{};
    sb.out_step = out_step;
    int first = 1;
    auto flush = [&]() {
        sb.accumulate = first ? 0 : 1;
        hipLaunchKernelGGL(sum_kernel, dim3(gb, n_rep), dim3(256), 0, g_->stream(), sb, M_, out);
        FTS_HIP_CHECK(hipGetLastError());
        first = 0;
        sb.n = 0;
    };
    for (int t = 0; t < n; t++) {
        const double* src;
        size_t step;
        if (kind[t] == FTS_SRC_BLOCK) {
            FTS_REQUIRE(idx[t] >= 0 && idx[t] + (size_t)(n_rep - 1) * block_step < blocks_.size(),
                        "block index out of range");
            src = phib_.as<double>() + (size_t)idx[t] * M_;
            step = (size_t)block_step * M_;
        } else {
            FTS_REQUIRE(idx[t] >= 0 && idx[t] + (size_t)(n_rep - 1) * solvent_step < svf_.size(),
                        "solvent index out of range");
            src = phis_.as<double>() + (size_t)idx[t] * M_;
            step = (size_t)solvent_step * M_;
        }
        sb.src[sb.n] = src;
        sb.scale[sb.n] = scale ? scale[t] : 1.0;
        sb.step[sb.n] = step;
        sb.n++;
        if (sb.n == MAX_BATCH) flush();
    }
    if (sb.n || first) flush();
}

template <typename T>
void Solver<T>::sum_concentrations(int n, const int* kind, const int* idx, const double* scale, double* out,
                                   bool on_device) {
    double* dst = on_device ? out : g_->scratch64.as<double>();
    sum_concentrations_replicas(n, kind, idx, scale, dst, 1, 0, 0, 0);
    if (!on_device) {
        FTS_HIP_CHECK(hipMemcpyAsync(out, dst, sizeof(double) * M_, hipMemcpyDeviceToHost, g_->stream()));
        FTS_HIP_CHECK(hipStreamSynchronize(g_->stream()));
    }
}

template <typename T>
void Solver<T>::propagator(int prop, int n, int hf, double* out, bool on_device) {
    FTS_REQUIRE(prop >= 0 && prop < (int)props_.size(), "propagator index out of range");
    FTS_REQUIRE(n >= 0 && n <= props_[prop].max_n, "segment index out of range");
    const T* src = slice_ptr(prop, n, hf);
    double* dst = on_device ? out : g_->scratch64.as<double>();
    hipLaunchKernelGGL((widen_kernel<T>), dim3(std::min(1024, ceil_div(M_, 256))), dim3(256), 0, g_->stream(), src,
                       dst, M_);
    FTS_HIP_CHECK(hipGetLastError());
    if (!on_device) FTS_HIP_CHECK(hipMemcpyAsync(out, dst, sizeof(double) * M_, hipMemcpyDeviceToHost, g_->stream()));
    FTS_HIP_CHECK(hipStreamSynchronize(g_->stream()));
}

}  // namespace

std::unique_ptr<SolverBase> make_solver(Grid* g, const fts_plan* plan, int precision, const int* bc,
                                        bool reduce_memory) {
    std::unique_ptr<SolverBase> s;
    if (precision == FTS_FP64) {
        s.reset(new Solver<double>(g, plan, bc, reduce_memory));
    } else if (precision == FTS_FP32) {
        FTS_REQUIRE(!bc, "the real-space method runs in FP64");
        FTS_REQUIRE(!g->big, "FP32 propagators need grid sizes whose prime factors are 2, 3 and 5");
        s.reset(new Solver<float>(g, plan, nullptr, reduce_memory));
    } else {
        throw Error("precision must be FTS_FP64 or FTS_FP32");
    }
    const char* graphs = std::getenv("FTS_HIP_GRAPHS");
    if (graphs && graphs[0] == '1') s->set_graphs(true);
    return s;
}

}  // namespace fts
