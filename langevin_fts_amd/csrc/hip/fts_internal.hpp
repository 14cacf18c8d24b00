// Internal classes behind the C ABI of include/fts_hip.h.
#pragma once

#include <array>
#include <cstdlib>
#include <memory>
#include <vector>

#include "../../../include/fts_hip.h"
#include "fts_common.hpp"
#include "fts_fft.hpp"
#include "fts_fft_ct.hpp"
#include "fts_fft_wl.hpp"

namespace fts {

// Opt-in timing of the propagator kernels with HIP events on the context's stream.
// Each timed launch records an event pair and its algorithmic bytes (the bytes the
// launch must read and write once: its fields, spectra and w, not re-reads).
class KernelTimer {
public:
    ~KernelTimer();
    void enable(int sample_every);
    bool active() const { return every_ != 0; }
    // a start event when this launch of the family is sampled, else nullptr
    hipEvent_t begin(hipStream_t st, int family);
    void end(hipStream_t st, hipEvent_t start, int family, double bytes);
    fts_kernel_stats get(int family);

private:
    struct Pending {
        hipEvent_t a, b;
        int family;
        double bytes;
    };
    hipEvent_t take();
    void drain(size_t n);
    int every_ = 0;
    long long count_[FTS_KERNEL_FAMILIES] = {};
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> pool_;
    fts_kernel_stats stats_[FTS_KERNEL_FAMILIES] = {};
};

}  // namespace fts

struct fts_context {
    int device = 0;
    hipStream_t stream = nullptr;
    fts::KernelTimer timer;
};

namespace fts {

constexpr int RED_BLOCKS = 256;   // first-stage workgroups of the deterministic reductions
constexpr int RED_THREADS = 256;

// Batched weighted inner product: out[item] = scale * sum_i dv_i * a_i * b_i * exp(wcoef * w_i)
// (a or b null: 1; w null: 1), dv_i = dV * mask_i.
template <typename T>
struct DotItem {
    const T* a;
    const T* b;
    const T* w;
    T wcoef;
    double scale;
};

template <typename T>
struct DotBatch {
    int n;
    DotItem<T> it[MAX_BATCH];
};

// batches travel as kernel arguments
static_assert(sizeof(SlabBatch<double>) <= 4000, "slab batch exceeds the kernel-argument budget");
static_assert(sizeof(XBatch<double>) <= 4000, "pencil batch exceeds the kernel-argument budget");
static_assert(sizeof(DotBatch<double>) <= 4000, "dot batch exceeds the kernel-argument budget");

template <typename T>
struct StressBatch;
struct StressGeom;

class Grid {
public:
    Grid(fts_context* ctx, int dim, const int* nx, const double* lx, const double* mask);
    void set_lx(const double* lx);

    fts_context* ctx;
    int dim;
    int nx[3];
    double lx[3];
    double dx[3];
    int tnx[3];     // 3D-mapped sizes (Pseudo.cpp:31-45)
    double tdx[3];
    int M, nzc, Mc;
    // row pitch and per-field size (complex elements) of the half spectra in device
    // buffers: nzc and Mc, or the padded pitch of the wave-local kernels
    int spitch = 0;     // FP64 spectrum row pitch (complex elements)
    int spitch32 = 0;   // FP32 spectrum row pitch
    size_t Ms = 0;      // spectrum slot (complex elements), for either precision
    double volume;
    double dV;
    std::vector<double> mask_h;
    DeviceBuffer mask64, mask32;
    AxisPlan ax, ay, az;
    DeviceBuffer tw64[3], tw32[3];
    SlabGeom sg;
    XGeom xg;
    int slab_threads = 0, x_threads = 0;
    int slab_maxr = 8;   // real values per thread in the slab kernel (8 or 16)
    bool big = false;        // some axis needs a radix outside {2, 3, 4, 5, 8}
    int ct_n = 0;        // N of the compile-time specialised kernels for an N^3 grid, 0: runtime kernels
    bool wave_local = false; // N^3 grids: one wavefront per line transform
    // launch shapes of the wave-local kernels per precision ([0] FP64, [1] FP32):
    // slab workgroup size and pencil kz columns per wavefront, with how they were chosen
    int slab_bd[2] = {0, 0};
    // per slab mode (forward, half step, Richardson step): a workgroup size that
    // replaces slab_bd for that mode, 0: slab_bd
    int slab_bd_mode[2][3] = {{0, 0, 0}, {0, 0, 0}};
    int pencil_nl[2] = {0, 0};
    std::string fft_schedule[2];
    template <typename T>
    void tune_fft_schedule();
    void set_fft_schedule(int precision, int slab_threads, int pencil_columns);
    DeviceBuffer partials;   // reduction scratch: MAX_BATCH * RED_BLOCKS doubles
    DeviceBuffer scalars;    // reduction results: MAX_BATCH doubles
    DeviceBuffer scratch64;  // general FP64 scratch of 2 fields

    std::string fft_error;   // why this grid has no FFT plans ("" when it has them)
    void setup_fft();
    void require_fft() const {
        if (!fft_error.empty()) throw Error(fft_error);
    }

    hipStream_t stream() const { return ctx->stream; }
    template <typename T>
    int spec_pitch() const {
        return sizeof(T) == sizeof(float) ? spitch32 : spitch;
    }
    template <typename T>
    const T* mask() const;
    template <typename T>
    const typename vec2<T>::type* tw(int axis) const;
    template <typename T>
    size_t slab_lds() const {
        return (size_t)(sg.rows * sg.P + sg.ny + sg.nz) * sizeof(typename vec2<T>::type);
    }
    template <typename T>
    size_t x_lds() const {
        return (size_t)(xg.nx * xg.P + xg.nx) * sizeof(typename vec2<T>::type);
    }
    template <typename T>
    void launch_slab(int mode, const SlabBatch<T>& b) const;
    template <typename T>
    void launch_xpass(const XBatch<T>& b, const T* bx, const T* by, const T* bz) const;
    // The stress of b.n terms from the slab pass's output spectra, with their x transforms
    // fused into the reduction (wave-local grids only): 3 partial sums per workgroup into
    // partials, term-major; returns the workgroups per term.
    template <typename T>
    int launch_stress_pencil(const StressBatch<T>& b, const StressGeom& sg, const T* bx, const T* by, const T* bz,
                             double* partials) const;
    // an upper bound of the workgroups per term of launch_stress_pencil (its tiles are 8 or 16 wide)
    int stress_pencil_blocks_max() const { return ct_n * ((ct_n / 2 + 1 + 7) / 8); }
    // first stage into partials, second stage into out (device, b.n doubles)
    template <typename T>
    void launch_dots(const DotBatch<T>& b, double* out) const;
    double inner_product(const double* a, const double* b, bool on_device);
    double integral(const double* a, bool on_device);
    template <typename T>
    void fft_forward(const T* in, typename vec2<T>::type* out, int batch);
    template <typename T>
    void fft_backward(const typename vec2<T>::type* in, T* out, int batch);
    // fft_backward through a caller's workspace of ws_fields spectra (Ms elements
    // each): enqueued on the stream, no allocation, no synchronisation
    template <typename T>
    void fft_backward_ws(const typename vec2<T>::type* in, T* out, int batch, typename vec2<T>::type* ws,
                         int ws_fields);
    // fft_backward_ws on the per-device no-allocation workspace
    template <typename T>
    void fft_backward_noalloc(const typename vec2<T>::type* in, T* out, int batch);
};

// Looks up the device address of the no-allocation workspace for `device` (which must
// be current) once per device; called when a context is created.
void noalloc_scratch_init(int device);

// Boltzmann factors of one (monomer, bond kind), separable along the three axes.
// Row t of bx/by/bz holds exp(-b^2 ds/6 (2 pi k_d / L_d)^2) along axis d
// (Pseudo.cpp:17-82 split into its three factors); t = 2m: full bond of monomer m,
// t = 2m + 1: half bond (b^2/2).
struct BondTables {
    DeviceBuffer bx64, by64, bz64, bx32, by32, bz32;
    void build(const Grid& g, const std::vector<double>& bsq, double ds);
    template <typename T>
    const T* x() const;
    template <typename T>
    const T* y() const;
    template <typename T>
    const T* z() const;
};

class SolverBase {
public:
    virtual ~SolverBase() = default;
    virtual void update_laplacian() = 0;
    virtual void compute_propagators(const double* const* w, const double* const* q_init, bool on_device) = 0;
    virtual void compute_concentrations() = 0;
    virtual void compute_stress() = 0;
    virtual double partition(int p) = 0;
    virtual double solvent_partition(int s) = 0;
    virtual void dq_dl(int p, double* out) = 0;
    virtual void sum_concentrations(int n, const int* kind, const int* idx, const double* scale, double* out,
                                    bool on_device) = 0;
    // the same sums for n_rep replicas of a replicated plan at once: replica r reads
    // block idx + r block_step, solvent idx + r solvent_step, writes out + r out_step
    // (device output)
    virtual void sum_concentrations_replicas(int n, const int* kind, const int* idx, const double* scale,
                                             double* out, int n_rep, int block_step, int solvent_step,
                                             size_t out_step) = 0;
    virtual void propagator(int prop, int n, int half, double* out, bool on_device) = 0;
    // bytes of device memory and of pinned host memory the solver holds
    virtual void memory_usage(size_t* device, size_t* host) const = 0;
    // replay compute_propagators + compute_concentrations from a captured HIP graph when
    // the inputs are device fields at the same addresses as the captured call
    virtual void set_graphs(bool on) = 0;
    virtual bool graphs_on() const = 0;
    // a caller replayed this solver's launches from a graph of its own: host-side caches
    // of the results (the partition functions) are stale
    virtual void results_changed() = 0;
    // what the captured launches depend on besides buffer addresses: box volume, FFT
    // launch shapes, version of the Boltzmann tables
    virtual std::vector<double> launch_signature() const = 0;
};

// bc: NULL for pseudo-spectral steps, else 2 dim FTS_BC_* codes (real-space steps, FP64 continuous)
// reduce_memory: propagator slices in pinned host memory (CudaComputationReduceMemory*.cu)
std::unique_ptr<SolverBase> make_solver(Grid* g, const fts_plan* plan, int precision, const int* bc = nullptr,
                                        bool reduce_memory = false);

constexpr int AM_MAX_HIST = 64;

// One mixing's step in a batched launch (an entry of a device table): what
// AndersonMixingDevice::calculate_new_fields decided on the host, with the device
// rows it reads and writes.
struct AmBatchItem {
    int mode;      // 0: no step, 1: simple mixing, 2: Anderson mixing
    int na;        // -1: input not in the history; else dot products d . dh[n] for n = 0..na
    int L;         // ring length (row stride of the dot-product ring)
    double mix;
    const double* w;   // input fields
    const double* d;   // input derivatives
    double* out;       // new fields
    double* dots_out;  // dot-product row receiving d . dh[n]
    const double* dots;   // this mixing's ring of dot-product rows (L x L)
    double* coef;         // na coefficients of the Gram solve
    int drows[AM_MAX_HIST];            // dot-product ring rows of history n = 0 (newest) .. na
    const double* wh[AM_MAX_HIST];     // field history rows n = 0 .. na
    const double* dh[AM_MAX_HIST];     // derivative history rows n = 0 .. na
};

// Host and device tables and reduction scratch of batched mixing steps
struct AmBatchWorkspace {
    DeviceBuffer host_table, table, partials;
    hipEvent_t copied = nullptr;   // recorded after the last copy of the host table
    int capacity = 0;
    void reserve(int n);
    AmBatchWorkspace() = default;
    AmBatchWorkspace(const AmBatchWorkspace&) = delete;
    AmBatchWorkspace& operator=(const AmBatchWorkspace&) = delete;
    ~AmBatchWorkspace() {
        if (copied) (void)hipEventDestroy(copied);
    }
};

class AndersonMixingDevice {
public:
    AndersonMixingDevice(fts_context* ctx, int n_var, int max_hist, double start_error, double mix_min,
                         double mix_init, bool reduce_memory = false);
    void reset();
    // staged: w and d are already the ring slots of staged_w() / staged_d() (the caller
    // wrote them there, which it may do when will_insert(err))
    void calculate_new_fields(double* w_new, const double* w, const double* d, double old_err, double err,
                              bool on_device, bool staged = false);
    // whether a call with this error level inserts into the history (CpuAndersonMixing.cpp:86-93)
    bool will_insert(double err) const { return err < start_error_ || n_anderson_ >= 0; }
    bool reduce_memory() const { return w_hist_.pinned_host(); }
    // the ring slots the next insertion fills (device, n_var doubles each)
    double* staged_w() const;
    double* staged_d() const;

    // Steps of several mixings of one n_var in one set of launches. For each r with
    // active[r], the host decisions and device results of
    // ams[r]->calculate_new_fields(out[r], w[r], d[r], old_err[r], err[r], true, staged[r]),
    // where an input the mixing keeps must be staged (already in its ring slots).
    static void calculate_batch(fts_context* ctx, const std::vector<AndersonMixingDevice*>& ams,
                                const std::vector<char>& active, const std::vector<double*>& out,
                                const std::vector<const double*>& w, const std::vector<const double*>& d,
                                const std::vector<double>& old_err, const std::vector<double>& err,
                                const std::vector<char>& staged, AmBatchWorkspace& ws);

    // the device work of a batched step from a device table of n items: dot products for
    // n = 0..max_na (items with a smaller na skip the rest), Gram solves of up to
    // max_solve coefficients, mixes; partials: n x AM_MAX_HIST x RED_BLOCKS doubles
    static void launch_batch(hipStream_t st, const AmBatchItem* dev, int n, int n_var, int max_na, int max_solve,
                             int L, double* partials);
    // what a device-resident planner of the same steps needs
    int n_var() const { return n_var_; }
    int max_hist() const { return max_hist_; }
    double start_error() const { return start_error_; }
    double mix_min() const { return mix_min_; }
    double mix_init() const { return mix_init_; }
    double* w_hist() const { return w_hist_.as<double>(); }
    double* d_hist() const { return d_hist_.as<double>(); }
    double* dots_ring() const { return dots_ring_.as<double>(); }
    double* coef() const { return coef_.as<double>(); }

private:
    // the host part of a step: mixing decisions, ring insertion and the rows the
    // device work reads (w, d: the inputs as they sit on the device)
    AmBatchItem plan(const double* w, const double* d, double* out, double old_err, double err);
    fts_context* ctx_;
    int n_var_, max_hist_, n_anderson_;
    double start_error_, mix_min_, mix_, mix_init_;
    // device rings of max_hist+1 rows: fields, derivatives and dot-product rows
    // (CircularBuffer.cpp semantics; the host keeps the indices)
    DeviceBuffer w_hist_, d_hist_, in_w_, in_d_, out_, partials_, dots_ring_, coef_;
    int ring_start_ = 0, ring_n_ = 0;
    int dots_start_ = 0, dots_n_ = 0;
};

class Cnn;

class SaddlePoint {
public:
    // one Anderson mixing per replica; n_blocks / n_solvents: per replica (replicated plans)
    SaddlePoint(fts_context* ctx, SolverBase* solver, std::vector<AndersonMixingDevice*> ams, int M,
                const fts_field_theory& ft, int n_blocks = 0, int n_solvents = 0);
    ~SaddlePoint();
    SaddlePoint(const SaddlePoint&) = delete;
    SaddlePoint& operator=(const SaddlePoint&) = delete;
    // replicas: run every saddle iteration and the Langevin update from captured HIP graphs
    // (the mixing decisions made on the device); off: the same launches issued one by one
    void set_graphs(bool on) { rep_graphs_ = on; }
    bool graphs() const { return rep_graphs_; }
    void set_fields(const double* w, bool on_device);
    void get_fields(double* w, bool on_device);
    void get_concentration(int type, double* out, bool on_device);
    void set_noise(const double* noise, bool on_device);
    void get_noise(double* noise, bool on_device);
    void backup();
    void restore(const int* replicas = nullptr);
    // n_iter, error and hamiltonian hold one entry per replica
    void find(double tolerance, int max_iter, int verbose, double* log, int* n_iter, double* error,
              double* hamiltonian);
    void langevin_update(const double* noise, bool on_device, double dt);
    void field_moments(double* mean, double* mean_sq);
    int n_replicas() const { return NR_; }
    // iterations run on `lo` (same plan, lower precision) while any active replica's
    // error is at or above switch_error; convergence is only declared on the main solver
    void set_low_precision(SolverBase* lo, double switch_error) {
        lo_ = lo;
        lo_switch_ = switch_error;
    }
    // learned guess of the imaginary field from the real one (single replica)
    void prepare_guess(Cnn& cnn);               // f(w_real) before the Langevin update (delta mode)
    // mode 0: w_imag = f(w_real); 1: w_imag += f(w_real) - prepared
    void predict_guess(Cnn& cnn, int mode);
    // residual guess (three-input network g), armed for the next search only: it keeps the
    // residual of its first evaluation (record), and with cnn set replaces the first
    // iteration's mixing by w_imag += g(w_real, first residual, w_imag)
    void set_residual_guess(Cnn* cnn, bool record);
    // the current training sample of the residual guess (4 M doubles: inputs, target), host
    void guess_sample(double* out);
    double train_guess(Cnn& cnn, double lr, int iters, int replay);

private:
    double* wa(int r) const { return w_aux_.as<double>() + (size_t)r * S_ * M_; }
    double* ph(int r) const { return phi_.as<double>() + (size_t)r * S_ * M_; }
    void concentrations(SolverBase* solver);
    void run_lincomb(int rep, const std::vector<double>& cw, const std::vector<double>& cp,
                     const std::vector<double>& c0, int n_out, double* out, int n_rep = 1, size_t out_step = 0);
    void row_sums(const std::vector<const double*>& rows, const double* centre, bool square, double* out_dev);
    void row_sums_strided(const double* base, int n, size_t stride, const double* centre, bool square,
                          double* out_dev);
    double hamiltonian(int rep, SolverBase* solver);
    void hamiltonians(const std::vector<int>& reps, SolverBase* solver, std::vector<double>& ham);
    void mix_replicas(const std::vector<char>& mix, const std::vector<double>& old_err,
                      const std::vector<double>& err, bool contiguous);
    // graph form of the replica iteration: A = evaluation, statistics and the device-side
    // mixing plan (one graph per precision), B = gather and batched Anderson step
    bool use_replica_graphs() const;
    void setup_replica_graphs(bool contiguous);
    void find_replicas_graph(double tolerance, int max_iter, bool contiguous, std::vector<int>& done,
                             std::vector<double>& err, std::vector<double>& ham);
    void launch_graph_a(SolverBase* solver, bool low, double tolerance, bool contiguous);
    void launch_graph_b(bool contiguous);
    void issue_graph_a(SolverBase* solver, bool low, double tolerance, bool contiguous);
    void issue_graph_b(bool contiguous);
    void issue_langevin(double dt);
    template <typename F>
    void capture(hipGraphExec_t& exec, F&& issue);
    bool rep_graphs_ = true;
    struct RepGraph {
        hipGraphExec_t exec = nullptr;
        const SolverBase* solver = nullptr;
        std::vector<double> sig;
        double tol = 0.0;
    };
    RepGraph rep_a_[2];                 // [0] main solver, [1] low-precision solver
    hipGraphExec_t rep_b_ = nullptr;
    struct LmGraph {                    // Langevin update, one per order of the two noise buffers
        hipGraphExec_t exec = nullptr;
        const void* prev = nullptr;
        double dt = 0.0;
    };
    LmGraph rep_lm_[2];
    // device-side mixing state, constants, step table and status (errors, done iterations)
    DeviceBuffer rep_state_, rep_const_, rep_items_, rep_status_, rep_status_host_;

    fts_context* ctx_;
    SolverBase* solver_;
    SolverBase* lo_ = nullptr;
    double lo_switch_ = 0.0;
    std::vector<AndersonMixingDevice*> ams_;
    int M_, NR_, n_blocks_, n_solvents_, S_, R_, I_, P_, n_solver_w_;
    std::vector<int> real_idx_, imag_idx_;
    std::vector<double> w_map_;   // n_solver_w x S
    std::vector<int> term_off_, term_kind_, term_index_;
    std::vector<double> term_scale_;
    std::vector<double> lam_cw_, lam_cp_, lam_c0_;   // R x S, R x S, R
    std::vector<double> h_cw_, h_cp_, h_c0_;         // I x S, I x S, I
    std::vector<double> dt_scaling_, h_lin_, h_sq_, pol_weight_;
    double h_const_;
    DeviceBuffer w_aux_, phi_, w_solver_, h_, lam_, am_w_, am_d_, noise_prev_, noise_cur_, w_backup_, phi_backup_,
        partials_, stats_;
    // replicas: per-replica gather destinations (pinned host table and its device copy)
    // and the batched mixing's tables
    DeviceBuffer gather_host_, gather_dev_;
    AmBatchWorkspace am_ws_;
    // learned guess: replay ring of (w_real, w_imag) pairs and the prepared prediction
    // the three inputs of the residual guess (and its target): the w+ the first residual
    // was evaluated at is the backup's (from_backup) or the current one
    void residual_inputs(double* in, double* target, bool from_backup);
    void predict_residual_correction(Cnn& cnn);   // the network's correction of w+ into guess_new_
    void apply_residual_guess(Cnn& cnn);
    DeviceBuffer replay_, guess_old_, guess_new_, guess_in_, res_first_, res_w0_;
    Cnn* res_cnn_ = nullptr;
    bool res_record_ = false, has_first_res_ = false;
    bool has_backup_ = false;
    int replay_per_ = 0;   // fields per ring sample: the network's inputs and its target
    // which network (and parameter version) and which real field a guess buffer holds f of:
    // the real field's version changes with every update of it and returns with a restore
    struct GuessTag {
        const Cnn* cnn = nullptr;
        long long cnn_version = -1, field_version = -1;
        bool operator==(const GuessTag& o) const {
            return cnn == o.cnn && cnn_version == o.cnn_version && field_version == o.field_version;
        }
    };
    GuessTag old_tag_, new_tag_;
    long long field_version_ = 0, backup_field_version_ = 0, version_counter_ = 0;
    int replay_cap_ = 0, replay_n_ = 0, replay_head_ = 0;
    unsigned long long replay_rng_ = 0x9E3779B97F4A7C15ull;
    bool prepared_ = false;
};

int cnn_n_params(int n_in);

// Periodic 3D CNN of the learned saddle-point guess (fts_cnn.hip)
class Cnn {
public:
    // n_in input channels, in_scales[n_in]: inputs are divided by their scale, outputs
    // multiplied by out_scale
    Cnn(fts_context* ctx, int nx, int ny, int nz, int n_in, const double* in_scales, double out_scale);
    // changes whenever the parameters do
    long long version() const { return version_; }
    void set_params(const float* host);
    void get_params(float* host);
    void get_grads(float* host);
    void predict(const double* w_minus, double* w_plus, bool on_device);
    // loss = mean (f(w-) - w+)^2 in scaled units; gradients into the device grads
    double gradients(const double* w_minus, const double* w_plus_target, bool on_device);
    double train_step(const double* w_minus, const double* w_plus_target, bool on_device, double lr);
    int n_grid() const { return M_; }
    int n_in() const { return n_in_; }
    int n_params() const { return n_params_; }

private:
    void load_input(const double* w_minus, bool on_device);
    // train: also write the channel-major copies of a[0..NHID-1] the weight gradients read
    void forward(bool train = false);
    void refresh_weights();
    fts_context* ctx_;
    int nx_, ny_, nz_, n_in_, n_params_, M_, n_chunks_;
    // hidden weight gradients from channel-major operands ([C][M] copies of the layer
    // inputs and of dZ) when nz % 8 == 0, over chunks of n_hchunks_
    bool cm_ = false;
    bool tiled_ = false;   // hidden layers through the LDS-tiled kernel (nz % 32 == 0, nz <= 128, ny % 4 == 0)
    void conv_hidden_launch(const void* X, const void* W, const float* bias, const void* mask, int relu, void* Y,
                            void* Ycm);
    int n_hchunks_ = 0;
    DeviceBuffer acts_cm_, gcm_a_, gcm_b_;
    std::vector<double> in_scales_;
    double out_scale_;
    long long step_ = 0;
    long long version_ = 0;
    DeviceBuffer params_, grads_, adam_m_, adam_v_, wf_, wb_, x_, target_, y_, dy_, acts_, grad_a_, grad_b_, partial_,
        loss_partials_, loss_, stage_in_, stage_out_;
};

}  // namespace fts

struct fts_grid {
    std::unique_ptr<fts::Grid> g;
};
struct fts_solver {
    std::unique_ptr<fts::SolverBase> s;
};
struct fts_am {
    std::unique_ptr<fts::AndersonMixingDevice> a;
};
struct fts_saddle {
    std::vector<std::unique_ptr<fts::AndersonMixingDevice>> owned_ams;   // replica saddles own theirs
    std::unique_ptr<fts::SaddlePoint> s;
};
struct fts_cnn {
    std::unique_ptr<fts::Cnn> c;
};
