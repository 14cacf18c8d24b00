/* C ABI of libfts_hip.so: the MI355X (gfx950) chain-propagator and concentration path.
 *
 * Plain pointers, sizes and opaque handles; no C++ or torch types. Every entry
 * point returns FTS_OK (0) or FTS_ERROR (nonzero) and leaves a message for
 * fts_last_error() (thread-local). Device pointers are HIP device addresses on the
 * context's device; "on_device" flags say whether a double* argument is a device
 * or a host address. All work is enqueued on the context's stream; calls that
 * return host values synchronise that stream.
 *
 * Each entry point names the reference interface it stands behind.
 */
#ifndef FTS_HIP_H
#define FTS_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fts_context fts_context;
typedef struct fts_grid fts_grid;
typedef struct fts_solver fts_solver;
typedef struct fts_am fts_am;
typedef struct fts_saddle fts_saddle;
typedef struct fts_cnn fts_cnn;

enum { FTS_OK = 0, FTS_ERROR = 1 };
enum { FTS_CONTINUOUS = 0, FTS_DISCRETE = 1 };
enum { FTS_FP64 = 0, FTS_FP32 = 1 };
/* kind of the initial slice of a propagator (PropagatorCode.cpp:117-247 key forms) */
enum { FTS_LEAF = 0, FTS_LEAF_QINIT = 1, FTS_JUNCTION = 2, FTS_AGGREGATED = 3 };
/* sources of fts_solver_sum_concentrations */
enum { FTS_SRC_BLOCK = 0, FTS_SRC_SOLVENT = 1 };
/* kernel families of fts_context_get_kernel_timing (FTS_KERNEL_REAL: the launches of
 * one real-space Crank-Nicolson step, timed together) */
enum { FTS_KERNEL_SLAB = 0, FTS_KERNEL_PENCIL = 1, FTS_KERNEL_REAL = 2, FTS_KERNEL_FAMILIES = 3 };
typedef struct fts_kernel_stats {
    double milliseconds;   /* summed launch durations (HIP events around each launch) */
    double bytes;          /* summed algorithmic bytes: each field, spectrum and w read or written once */
    long long launches;
} fts_kernel_stats;
/* hidden channels of the learned-guess network */
#define FTS_CNN_CHANNELS 32
/* most monomer types of a field theory handled by fts_saddle */
#define FTS_MAX_TYPES 8

/* Last error message of the calling thread ("" if none). */
const char* fts_last_error(void);
/* ABI version, bumped on incompatible changes. */
int fts_abi_version(void);

/* ---- context: one device and one stream (replaces CudaCommon's singleton,
 *      CudaCommon.cu:15-113) */
int fts_context_create(int device, fts_context** out);
int fts_context_destroy(fts_context* ctx);
/* Kernel timing (no reference counterpart: measurement of the slab and pencil
 * passes for roofline reports). sample_every = k > 0 times one launch in k of each
 * family with a pair of fence-free HIP events (0: off); the call clears the totals.
 * get synchronises the stream and returns the totals of the timed launches. */
int fts_context_set_kernel_timing(fts_context* ctx, int sample_every);
int fts_context_get_kernel_timing(fts_context* ctx, int family, fts_kernel_stats* out);
int fts_context_synchronize(fts_context* ctx);
/* hipStream_t of the context, returned as void* */
int fts_context_stream(fts_context* ctx, void** stream);
/* device description ("name, gcnArch, CUs, memory") into buf (AbstractFactory::display_info) */
int fts_context_info(fts_context* ctx, char* buf, int buflen);

/* ---- grid: ComputationBox (ComputationBox.h:22-76, CudaComputationBox.cu:17-118)
 * dim 1..3, nx and lx of length dim, mask NULL or nx-product values of 0/1. */
int fts_grid_create(fts_context* ctx, int dim, const int* nx, const double* lx, const double* mask, fts_grid** out);
int fts_grid_destroy(fts_grid* g);
/* ComputationBox::set_lx (ComputationBox.cpp:229-262) */
int fts_grid_set_lx(fts_grid* g, const double* lx);
int fts_grid_volume(fts_grid* g, double* out);
/* ComputationBox::integral / inner_product (ComputationBox.cpp:265-279) */
int fts_grid_integral(fts_grid* g, const double* a, int on_device, double* out);
int fts_grid_inner_product(fts_grid* g, const double* a, const double* b, int on_device, double* out);
/* Batched real 3D FFT on device buffers, the transform of MklFFT3D::forward /
 * backward (MklFFT3D.cpp:55-79): forward unnormalised real -> half complex
 * [nx][ny][nz/2+1] (interleaved re,im), backward half complex -> real divided by
 * the grid size. precision FTS_FP64 (double) or FTS_FP32 (float). In-place is not
 * allowed. batch fields are contiguous. */
int fts_grid_fft_forward(fts_grid* g, int precision, const void* d_in, void* d_out, int batch);
int fts_grid_fft_backward(fts_grid* g, int precision, const void* d_in, void* d_out, int batch);
/* No counterpart in the reference: fts_grid_fft_backward for callers that cannot
 * allocate device memory (code under HIP-graph capture, serving loops, harnesses
 * that forbid hipMalloc). Same arguments and results (bit for bit); it allocates
 * nothing and does not synchronise: the work is enqueued on the context's stream.
 * Its workspace is a fixed 32 MiB array per GPU, kept until the process ends: a grid
 * whose half spectrum (nx ny (nz/2+1) rounded up to rows of 8, times 16 bytes in
 * FP64, 8 in FP32) exceeds it is refused with an error (128^3 FP64 needs 18.9 MB).
 * Every call on one GPU shares that array, so calls whose work overlaps on one GPU
 * (other streams or contexts on the device) race; calls in one stream do not. */
int fts_grid_fft_backward_noalloc(fts_grid* g, int precision, const void* d_in, void* d_out, int batch);
/* FFT schedule of N^3 grids (N = 32, 48, 64, 96, 128); no counterpart in the
 * reference, whose FFTs are library calls (MklFFT3D.cpp:11-79, CudaFFT). Per
 * precision, the grid picks the slab-pass workgroup size (slab_threads) and the kz
 * columns per wavefront of the pencil pass (pencil_columns) by timing the candidates
 * when it is created. Results of the candidates are bit-identical at 32^3 and 64^3,
 * where tuning is on by default; at 48^3, 96^3 and 128^3 they agree to rounding only,
 * so the defaults stay there unless the environment sets FTS_FFT_TUNE=1 (a timed
 * choice could differ between processes and with it a chaotic run's trajectory).
 * FTS_FFT_TUNE=0 keeps the defaults everywhere. The default pencil_columns is 8, and
 * 4 in FP64 from 96^3. One process keeps one choice per (N, precision). slab_threads is
 * the half step's (discrete
 * chains); the forward transform with two spectra out and the Richardson step of
 * continuous chains are timed alone and may take another candidate, which the note
 * names. get writes the choice and a note with the measured candidates
 * (NUL-terminated, truncated to notelen); set pins one candidate for every slab mode.
 * Other grids report 0 and refuse set. */
int fts_grid_get_fft_schedule(fts_grid* g, int precision, int* slab_threads, int* pencil_columns, char* note,
                              int notelen);
int fts_grid_set_fft_schedule(fts_grid* g, int precision, int slab_threads, int pencil_columns);

/* ---- propagator computation plan: the output of PropagatorAnalyzer
 * (PropagatorAnalyzer.cpp:142-269) and Scheduler (Scheduler.cpp:11-190) as flat
 * arrays. Propagators are indexed in ComparePropagatorKey order, blocks in
 * std::map<(polymer, left key, right key)> order, monomers in name order. */
typedef struct {
    int model;                   /* FTS_CONTINUOUS or FTS_DISCRETE */
    double ds;
    int n_monomers;
    const double* segment_length; /* statistical segment length a_m; bond b^2 = a_m^2 */
    int n_props;
    const int* prop_monomer;
    const int* prop_max_n;       /* max_n_segment */
    const int* prop_kind;        /* FTS_LEAF ... FTS_AGGREGATED */
    const int* prop_qinit;       /* q_init index for FTS_LEAF_QINIT, else -1 */
    const int* prop_dep_off;     /* into dep_* */
    const int* prop_n_deps;
    const int* dep_prop;
    const int* dep_n;
    const int* dep_rep;
    const int* prop_je_off;      /* into je (discrete junction_ends) */
    const int* prop_n_je;
    const int* je;
    int n_blocks;
    const int* blk_polymer;
    const int* blk_left;
    const int* blk_right;
    const int* blk_n_left;
    const int* blk_n_right;
    const int* blk_rep;
    const int* blk_monomer;
    const int* blk_n_vu;
    int n_polymers;
    const double* pol_vf;
    const double* pol_alpha;
    int n_solvents;
    const double* sol_vf;
    const int* sol_monomer;
    int n_qinit;
    int n_slices;                /* schedule: slice s holds jobs slice_off[s] .. slice_off[s+1]-1 */
    const int* slice_off;
    const int* job_prop;
    const int* job_from;
    const int* job_to;
} fts_plan;

/* CpuComputationContinuous / CpuComputationDiscrete constructors
 * (CpuComputationContinuous.cpp:9-105, CpuComputationDiscrete.cpp:8-137);
 * precision FTS_FP64 or FTS_FP32 (propagators and FFTs; Q, concentrations and
 * stress accumulate in FP64 either way). The plan arrays are copied. */
int fts_solver_create(fts_grid* g, const fts_plan* plan, int precision, fts_solver** out);
/* boundary conditions of fts_solver_create_realspace (ComputationBox.h:15-20) */
enum { FTS_BC_PERIODIC = 0, FTS_BC_REFLECTING = 1, FTS_BC_ABSORBING = 2 };
/* CpuComputationContinuous constructed with method "realspace" (CpuComputationContinuous.cpp:25-26,
 * CpuSolverReal.cpp:9-45; AbstractFactory::create_realspace_solver, AbstractFactory.h:53):
 * continuous chains only, FP64, one Crank-Nicolson ADI step per contour step.
 * bc: 2 dim codes FTS_BC_*, (low, high) per axis in axis order; both sides of an
 * axis are periodic or neither is. fts_solver_compute_stress fails on such a solver,
 * as CpuSolverReal::compute_single_segment_stress_continuous does. */
int fts_solver_create_realspace(fts_grid* g, const fts_plan* plan, const int* bc, fts_solver** out);
/* The reduce-memory forms CudaComputationReduceMemoryContinuous / ...Discrete
 * (CudaComputationReduceMemoryContinuous.cu:13-120, CudaComputationReduceMemoryDiscrete.cu; selected by
 * PlatformSelector::create_factory(name, reduce_memory_usage = true), PlatformSelector.cpp:16-54):
 * every propagator slice lives in pinned host memory, which the kernels write and read in
 * place over the host link; fields, spectra and concentrations stay in device memory.
 * Results equal those of fts_solver_create / fts_solver_create_realspace.
 * bc: NULL for pseudo-spectral steps, else as fts_solver_create_realspace (then FP64). */
int fts_solver_create_reduce_memory(fts_grid* g, const fts_plan* plan, int precision, const int* bc,
                                    fts_solver** out);
/* Bytes of device memory and of pinned host memory a solver holds (no reference
 * counterpart: lets callers and tests see where the propagators live). */
int fts_solver_get_memory_usage(fts_solver* s, size_t* device_bytes, size_t* host_bytes);
/* n_replicas independent copies of the plan in one solver (no reference counterpart:
 * the replica-batched L-FTS form). Replica r has its own monomer fields (index
 * r n_monomers + m in the w array of compute_propagators), blocks (r n_blocks + b),
 * polymers (r n_polymers + p) and solvents; every schedule slice advances all
 * replicas' jobs, so the same launches carry all of them. */
int fts_solver_create_replicas(fts_grid* g, const fts_plan* plan, int precision, int n_replicas, fts_solver** out);
int fts_solver_destroy(fts_solver* s);
/* PropagatorComputation::update_laplacian_operator after fts_grid_set_lx */
int fts_solver_update_laplacian(fts_solver* s);
/* PropagatorComputation::compute_propagators (CpuComputationContinuous.cpp:148-369,
 * CpuComputationDiscrete.cpp:192-582). w: n_monomers field pointers, q_init:
 * n_qinit pointers (or NULL when n_qinit == 0), each of grid size, FP64. */
int fts_solver_compute_propagators(fts_solver* s, const double* const* w, const double* const* q_init,
                                   int on_device);
/* compute_concentrations (CpuComputationContinuous.cpp:371-443, CpuComputationDiscrete.cpp:584-663) */
int fts_solver_compute_concentrations(fts_solver* s);
/* HIP graphs of the evaluation; no counterpart in the reference (its CUDA platform
 * launches eagerly). enable != 0: compute_propagators with on_device != 0 is captured
 * once into a HIP graph and replayed while the field (and q_init) addresses, the box
 * volume and the FFT launch shapes are those of the captured call, else captured
 * again; compute_concentrations after it likewise. With kernel timing on, calls run
 * eagerly. Results equal the eager calls'. Environment FTS_HIP_GRAPHS=1 sets it for
 * every solver at creation. */
int fts_solver_set_graphs(fts_solver* s, int enable);
/* compute_stress (CpuComputationContinuous.cpp:643-726, CpuComputationDiscrete.cpp:857-996) */
int fts_solver_compute_stress(fts_solver* s);
/* get_total_partition / get_solvent_partition */
int fts_solver_get_partition(fts_solver* s, int polymer, double* out);
int fts_solver_get_solvent_partition(fts_solver* s, int solvent, double* out);
/* dq_dl[polymer][d] for d < dim (before the phi/alpha/Q normalisation of
 * PropagatorComputation::get_stress, PropagatorComputation.cpp:35-52) */
int fts_solver_get_dq_dl(fts_solver* s, int polymer, double* out);
/* out = sum_t scale[t] * phi(src_kind[t], src_index[t]) over n_terms sources:
 * get_total_concentration(_gce), get_block_concentration, get_solvent_concentration */
int fts_solver_sum_concentrations(fts_solver* s, int n_terms, const int* src_kind, const int* src_index,
                                  const double* scale, double* out, int on_device);
/* get_chain_propagator: q[prop][n] (half = 1: discrete half-bond slice), FP64 out */
int fts_solver_get_propagator(fts_solver* s, int prop, int n, int half, double* out, int on_device);

/* ---- Anderson mixing (AndersonMixing.cpp:3-47, CpuAndersonMixing.cpp:76-161,
 *      CudaAndersonMixing.cu:9-312); history kept on the device, FP64 Gram solve */
/* CudaAndersonMixingReduceMemory (CudaAndersonMixingReduceMemory.cu): the history rings of
 * fts_am_create in pinned host memory; same results. */
int fts_am_create_reduce_memory(fts_context* ctx, int n_var, int max_hist, double start_error, double mix_min,
                                double mix_init, fts_am** out);
int fts_am_create(fts_context* ctx, int n_var, int max_hist, double start_error, double mix_min, double mix_init,
                  fts_am** out);
int fts_am_destroy(fts_am* am);
int fts_am_reset(fts_am* am);
int fts_am_calculate_new_fields(fts_am* am, double* w_new, const double* w_current, const double* w_deriv,
                                double old_error_level, double error_level, int on_device);

/* ---- L-FTS saddle point and Langevin update, device-resident
 *      (examples/fts/lfts.py:631-868: LFTS.run, find_saddle_point,
 *      compute_concentrations; SymmetricPolymerTheory.compute_func_deriv and
 *      compute_hamiltonian, lfts.py:197-245). The field theory arrives as the
 *      small coefficient tables the driver computes on the host:
 *        solver field m   = sum_j w_map[m][j] w_aux[j]                (lfts.py:534-545)
 *        phi[type i]      = sum_t scale_t * (block or solvent term t)  (lfts.py:550-565)
 *        dH/dw_aux[imag c] = sum_j imag_cw[c][j] w_aux[j] + imag_cp[c][j] phi[j] + imag_c0[c]
 *                           (sign of lfts.py:243-244 already applied)
 *        dH/dw_aux[real c] = the same with real_cw, real_cp, real_c0
 *        H = sum_j h_lin[j] <w_aux[j]> + h_sq[j] <w_aux[j]^2>
 *            - sum_p polymer_weight[p] ln Q_p + h_const               (lfts.py:197-216)
 *      <.> is the plain grid mean (numpy mean). */
typedef struct fts_field_theory {
    int n_types;           /* S, at most FTS_MAX_TYPES; w_aux and phi are [S][n_grid] */
    int n_real, n_imag;    /* R and I */
    const int* real_idx;   /* R indices into w_aux */
    const int* imag_idx;   /* I indices into w_aux (the last one is the pressure field) */
    int n_solver_fields;   /* monomer fields the solver takes, in its monomer order (<= 8) */
    const double* w_map;   /* n_solver_fields x S */
    const int* phi_term_off;      /* S + 1 offsets into the term arrays */
    const int* phi_term_kind;     /* FTS_SRC_BLOCK or FTS_SRC_SOLVENT */
    const int* phi_term_index;    /* block or solvent index of the solver */
    const double* phi_term_scale;
    const double* real_cw; const double* real_cp; const double* real_c0;   /* R x S, R x S, R */
    const double* imag_cw; const double* imag_cp; const double* imag_c0;   /* I x S, I x S, I */
    const double* dt_scaling;     /* S (lfts.py:498-501) */
    const double* h_lin;          /* S */
    const double* h_sq;           /* S */
    double h_const;
    int n_polymers;
    const double* polymer_weight; /* volume fraction / alpha per polymer */
} fts_field_theory;

int fts_saddle_create(fts_context* ctx, fts_solver* solver, fts_am* am, int n_grid, const fts_field_theory* ft,
                      fts_saddle** out);
/* n_replicas independent saddle points over a solver made by fts_solver_create_replicas
 * (n_blocks, n_solvents: per replica; one Anderson mixing per replica with the given
 * parameters). Fields, concentrations and noise then hold n_replicas consecutive
 * copies ([R][S][n_grid], [R][S][n_grid], [R][n_real][n_grid]); find iterates until
 * every replica converged or max_iter (a converged replica keeps its fields) and
 * writes n_iter, error_level and hamiltonian per replica; field_moments returns R x S
 * values. The learned-guess calls need a single replica. */
int fts_saddle_create_replicas(fts_context* ctx, fts_solver* solver, int n_replicas, int n_blocks, int n_solvents,
                               int max_hist, double start_error, double mix_min, double mix_init, int n_grid,
                               const fts_field_theory* ft, fts_saddle** out);
int fts_saddle_destroy(fts_saddle* s);
/* w_aux [S][n_grid] in and out */
int fts_saddle_set_fields(fts_saddle* s, const double* w_aux, int on_device);
int fts_saddle_get_fields(fts_saddle* s, double* w_aux, int on_device);
/* phi of monomer type i from the last saddle iteration */
int fts_saddle_get_concentration(fts_saddle* s, int type, double* out, int on_device);
/* normal_noise_prev [R][n_grid] (lfts.py:669-672, :712; checkpoint state) */
int fts_saddle_set_noise(fts_saddle* s, const double* noise_prev, int on_device);
int fts_saddle_get_noise(fts_saddle* s, double* noise_prev, int on_device);
/* copy of w_aux and phi taken before a Langevin step and put back when the saddle
 * point fails (lfts.py:696-698, :727-729) */
int fts_saddle_backup(fts_saddle* s);
int fts_saddle_restore(fts_saddle* s);
/* put back only the replicas whose flag is nonzero (n_replicas flags) */
int fts_saddle_restore_replicas(fts_saddle* s, const int* replicas);
/* find_saddle_point (lfts.py:806-867). verbose 2: log row per iteration; verbose 1:
 * one row for the final iteration; row = [iter, mass error, H, Q_0..Q_{P-1},
 * error_0..error_{I-1}] (3 + P + I doubles). hamiltonian: H at the final iteration. */
int fts_saddle_find(fts_saddle* s, double tolerance, int max_iter, int verbose, double* log, int* n_iter,
                    double* error_level, double* hamiltonian);
/* Leimkuhler-Matthews update of the real fields with this step's noise [R][n_grid]
 * (lfts.py:701-713); the noise then becomes normal_noise_prev. */
int fts_saddle_langevin_update(fts_saddle* s, const double* noise, int on_device, double dt);
/* plain means of w_aux[j] and w_aux[j]^2, S each (lfts.py:250-262 dH/dchiN terms) */
int fts_saddle_field_moments(fts_saddle* s, double* mean, double* mean_sq);

/* ---- Learned initial guess of the saddle point (no counterpart in the
 *      reference, whose drivers start each saddle point from the previous step's
 *      w+, examples/fts/lfts.py:701-717). A periodic 3D CNN of n_in input fields:
 *      conv(n_in->32) ReLU, 4 x [conv(32->32) ReLU], conv(32->1), 3x3x3 kernels with
 *      circular padding; bf16 weights and activations, FP32 accumulation (MFMA);
 *      FP32 master weights trained with Adam, laid out W0[n_in][27][32], b0[32],
 *      4 x (W[27][32][32], b[32]), WL[27][32], bL. Input channel c is divided by
 *      in_scales[c] and the output multiplied by out_scale; inputs are n_in
 *      consecutive fields. The grid size must be a multiple of 32.
 * n_params: the count of a one-input network; param_count: this network's count. */
int fts_cnn_n_params(void);
int fts_cnn_create(fts_context* ctx, const int* nx, double in_scale, double out_scale, fts_cnn** out);
int fts_cnn_create_inputs(fts_context* ctx, const int* nx, int n_in, const double* in_scales, double out_scale,
                          fts_cnn** out);
int fts_cnn_param_count(fts_cnn* c, int* n);
int fts_cnn_destroy(fts_cnn* c);
/* FP32 master parameters, fts_cnn_n_params() floats, host memory; setting resets Adam */
int fts_cnn_set_params(fts_cnn* c, const float* params);
int fts_cnn_get_params(fts_cnn* c, float* params);
int fts_cnn_predict(fts_cnn* c, const double* w_minus, double* w_plus, int on_device);
/* loss = grid mean of (f(w- / in_scale) - w+ / out_scale)^2; grads (host, n_params floats) may be null */
int fts_cnn_gradients(fts_cnn* c, const double* w_minus, const double* w_plus_target, int on_device, double* loss,
                      float* grads);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on one (w-, w+) pair */
int fts_cnn_train_step(fts_cnn* c, const double* w_minus, const double* w_plus_target, int on_device, double lr,
                       double* loss);
/* Mixed-precision saddle point (no reference counterpart; an addition for the
 * L-FTS throughput): while the error of any active replica is at or above
 * switch_error, the iterations evaluate the propagators with `lo`, a solver made
 * from the same plan and replica count at FTS_FP32; below it they use the
 * saddle's own solver, and only those evaluations can declare convergence, so the
 * returned fields, concentrations and Hamiltonian meet the tolerance in the main
 * solver's precision. lo = NULL switches it off. lo must outlive its use here. */
int fts_saddle_set_low_precision(fts_saddle* s, fts_solver* lo, double switch_error);
/* Replica saddles (no counterpart in the reference, whose replicas are separate runs):
 * with on != 0 (the default) every saddle iteration of fts_saddle_find and every
 * Langevin update replays captured HIP graphs. An iteration is two graphs: the
 * evaluation with the residual statistics and a planning kernel that makes every
 * replica's convergence and mixing decisions on the device, then the batched
 * Anderson step. The host reads the errors back once per iteration. Results equal
 * those of on = 0, which issues the same launches one by one. Graphs are recaptured
 * when the solver, its box, FFT launch shapes or Boltzmann tables, the tolerance or
 * the time step change; with kernel timing on, launches run one by one. No effect on
 * single-replica saddles. */
int fts_saddle_set_graphs(fts_saddle* s, int on);

/* Field theories with one real and one imaginary field (AB melts), on the device.
 * With a one-input network f(w-):
 *   prepare: remember f(w_aux[real]) before the Langevin update;
 *   predict: w_aux[imag] = f(w_aux[real]) (mode 0), or w_aux[imag] += f(w_aux[real]) - the
 *            prepared prediction (mode 1: the change of w+ the network expects from the update);
 *   train:   push the current (w_aux[real], w_aux[imag]) pair into a ring of `replay` samples,
 *            then `iters` Adam steps, the first on the newest sample and the rest on samples
 *            drawn from the ring; loss is the newest sample's loss before its step.
 * Residual guess, a three-input network g:
 *   set_residual_guess: arms the next fts_saddle_find (one search only): with record != 0
 *            it keeps the residual of its first evaluation (the pressure-field derivative at
 *            the starting w+); with c set, that first iteration's mixing is replaced by
 *            w_aux[imag] += g(w_aux[real], first residual, w_aux[imag]) and the Anderson
 *            mixing starts from there. The first evaluation counts as an iteration.
 *   train:   (a three-input network) pushes ([w_aux[real], first residual, w_aux[imag] of the
 *            last fts_saddle_backup], converged w_aux[imag] minus that backup's) into the ring;
 *   get_guess_sample: that sample (4 n_grid doubles, host) without training. */
int fts_saddle_set_residual_guess(fts_saddle* s, fts_cnn* c, int record);
int fts_saddle_get_guess_sample(fts_saddle* s, double* out);
int fts_saddle_prepare_guess(fts_saddle* s, fts_cnn* c);
int fts_saddle_predict_guess(fts_saddle* s, fts_cnn* c, int mode);
int fts_saddle_train_guess(fts_saddle* s, fts_cnn* c, double lr, int iters, int replay, double* loss);

#ifdef __cplusplus
}
#endif
#endif
