// Device-resident saddle-point iteration and Leimkuhler-Matthews update of the
// Langevin field-theoretic simulation (examples/fts/lfts.py:631-868).
//
// The auxiliary fields w_aux[S][M], the concentrations phi[S][M], the functional
// derivatives and both Langevin noise arrays stay in HBM for the whole run. One
// saddle iteration is: monomer fields as linear combinations of w_aux
// (lfts.py:163-164, :534-545), propagators + concentrations (the solver), per-type
// concentration sums, dH/dw of the imaginary fields (lfts.py:219-245), their
// standard deviations (lfts.py:829-831), Anderson mixing on the imaginary rows
// (lfts.py:860). Only 2 I doubles (mean and std of each derivative) and the
// Anderson dot products cross to the host per iteration.
#include <algorithm>
#include <cmath>
#include <vector>

#include "fts_internal.hpp"
#include "fts_reduce.hpp"

namespace fts {

namespace {

constexpr int LC_MAX_OUT = 8;
constexpr int LC_MAX_IN = 2 * FTS_MAX_TYPES;
constexpr int STAT_BLOCKS = 128;
constexpr int STAT_THREADS = 256;

// out[r][i] = c0[r] + sum_j c[r][j] * in[j][i]; blockIdx.y = replica q reads in[j] + q in_step
// and writes out[r] + q out_step
struct LinComb {
    int n_out, n_in;
    const double* in[LC_MAX_IN];
    double* out[LC_MAX_OUT];
    double c[LC_MAX_OUT][LC_MAX_IN];
    double c0[LC_MAX_OUT];
    size_t in_step, out_step;
};

__global__ void __launch_bounds__(256) lincomb_kernel(LinComb a, int M) {
    const size_t qi = blockIdx.y * a.in_step, qo = blockIdx.y * a.out_step;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        double x[LC_MAX_IN];
#pragma unroll
        for (int j = 0; j < LC_MAX_IN; j++)
            if (j < a.n_in) x[j] = a.in[j][qi + i];
#pragma unroll
        for (int r = 0; r < LC_MAX_OUT; r++) {
            if (r >= a.n_out) break;
            double v = a.c0[r];
#pragma unroll
            for (int j = 0; j < LC_MAX_IN; j++)
                if (j < a.n_in) v += a.c[r][j] * x[j];
            a.out[r][qo + i] = v;
        }
    }
}

// Row statistics of up to LC_MAX_OUT rows: sum, then sum of squared deviations from
// the row mean, each with a fixed-order two-stage reduction.
struct Rows {
    const double* row[2 * FTS_MAX_TYPES];
};

// this workgroup's share of sum_i (x_i - c) (or its square) over a grid of gridDim.x workgroups
__device__ __forceinline__ double row_share(const double* x, int M, double c, int square, double* ws) {
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        const double v = x[i] - c;
        acc += square ? v * v : v;
    }
    return block_sum(acc, ws);
}

__global__ void __launch_bounds__(STAT_THREADS) row_sum_stage1(Rows r, int M, const double* centre, int square,
                                                              double* partials) {
    __shared__ double ws[16];
    const double c = centre ? centre[blockIdx.y] / M : 0.0;
    const double acc = row_share(r.row[blockIdx.y], M, c, square, ws);
    if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = acc;
}

// rows base + k stride, k = blockIdx.y (the same per-row sums as row_sum_stage1)
__global__ void __launch_bounds__(STAT_THREADS) row_sum_strided_stage1(const double* base, size_t stride, int M,
                                                                      const double* centre, int square,
                                                                      double* partials) {
    __shared__ double ws[16];
    const double c = centre ? centre[blockIdx.y] / M : 0.0;
    const double acc = row_share(base + blockIdx.y * stride, M, c, square, ws);
    if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = acc;
}

__global__ void __launch_bounds__(STAT_BLOCKS) row_sum_stage2(const double* partials, double* out) {
    __shared__ double ws[16];
    double v = partials[blockIdx.x * blockDim.x + threadIdx.x];
    v = block_sum(v, ws);
    if (threadIdx.x == 0) out[blockIdx.x] = v;
}

// w_aux[S-1] -= mean(w_aux[S-1]) with the row sum already on the device; blockIdx.y =
// replica q: row x + q step, sum[q]
__global__ void subtract_mean_kernel(double* x, size_t step, const double* sum, int M) {
    const double m = sum[blockIdx.y] / M;
    x += blockIdx.y * step;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) x[i] -= m;
}

// Anderson input of the imaginary fields: w_im[c] = w_aux[idx[c]], d[c] = -scale[c] * h[c]
struct Gather {
    int n;
    const double* src[FTS_MAX_TYPES];
    const double* h[FTS_MAX_TYPES];
    double* w_out[FTS_MAX_TYPES];
    double* d_out[FTS_MAX_TYPES];
    double scale[FTS_MAX_TYPES];
};

__global__ void gather_kernel(Gather g, int M) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x)
        for (int c = 0; c < g.n; c++) {
            if (g.w_out[c]) g.w_out[c][i] = g.src[c][i];
            g.d_out[c][i] = -g.scale[c] * g.h[c][i];
        }
}

// the gather of several replicas at once: per replica its destinations (a table in
// device memory), the sources by stride from replica 0's
struct GatherRep {
    double* w_out[FTS_MAX_TYPES];   // null: the fields are read where they are
    double* d_out[FTS_MAX_TYPES];   // null: the replica is not mixed this iteration
    int scatter;                    // copy its mixed fields (am_w) back into w_aux
};

struct GatherShared {
    int n;
    double* wim[FTS_MAX_TYPES];         // imaginary w_aux rows of replica 0; replica q at + q wim_step
    const double* h[FTS_MAX_TYPES];     // derivative rows of replica 0; + q h_step
    double scale[FTS_MAX_TYPES];
    size_t wim_step, h_step;
    const double* am_w;                 // mixed fields of replica 0 (n rows); + q am_step
    size_t am_step;
};

__global__ void gather_batch_kernel(GatherShared g, const GatherRep* reps, int M) {
    const GatherRep& gr = reps[blockIdx.y];
    const size_t q = blockIdx.y;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x)
        for (int c = 0; c < g.n; c++) {
            if (gr.w_out[c]) gr.w_out[c][i] = g.wim[c][q * g.wim_step + i];
            if (gr.d_out[c]) gr.d_out[c][i] = -g.scale[c] * g.h[c][q * g.h_step + i];
        }
}

__global__ void scatter_batch_kernel(GatherShared g, const GatherRep* reps, int M) {
    if (!reps[blockIdx.y].scatter) return;
    const size_t q = blockIdx.y;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x)
        for (int c = 0; c < g.n; c++) g.wim[c][q * g.wim_step + i] = g.am_w[q * g.am_step + (size_t)c * M + i];
}

// ---- device-side mixing plan of the replica graphs: the host decisions of the
// unbatched loop (error level, convergence, AndersonMixingDevice::plan) per replica
struct RepAmState {
    int n_anderson, ring_start, ring_n, dots_start, dots_n;
    int done;   // iteration at which the replica converged, 0: not yet
    int it;     // iterations of this search
    int pad;
    double mix, err;
};

struct RepAmConst {   // the replica's mixing buffers
    double* w_hist;
    double* d_hist;
    double* dots;
    double* coef;
};

struct RepPlanArgs {
    int NR, I, M, L, max_hist, contiguous, low, nrows;
    double tol, start_error, mix_min, mix_init;
    double* am_w;   // non-staged inputs of replica 0; replica r at + r am_step
    double* am_d;
    size_t am_step;
    double* w_im;   // contiguous imaginary rows of replica 0; + r w_step
    size_t w_step;
};

__global__ void replica_reset_kernel(RepAmState* st, int NR, double mix_init) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= NR) return;
    RepAmState s{};
    s.n_anderson = -1;
    s.mix = mix_init;
    s.err = 1e20;
    st[r] = s;
}

// one thread per replica; stats: sums then centred sums of squares of the derivative rows
__global__ void replica_plan_kernel(RepPlanArgs a, const double* stats, RepAmState* st, const RepAmConst* cs,
                                    AmBatchItem* items, GatherRep* gather, double* status) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.NR) return;
    // the error level of lfts.py:829-831: the largest deviation of the imaginary derivatives
    double e = 0.0;
    for (int c = 0; c < a.I; c++) {
        const double ec = sqrt(stats[a.nrows + r * a.I + c] / a.M);
        if (c == 0 || ec > e || isnan(ec)) e = ec;
    }
    RepAmState s = st[r];
    AmBatchItem& it = items[r];
    it.mode = 0;
    it.na = -1;
    it.L = a.L;
    GatherRep g{};
    if (!s.done) {
        s.it += 1;
        const double old_err = s.err;
        s.err = e;
        if (!a.low && e < a.tol) {
            s.done = s.it;
        } else {
            const RepAmConst k = cs[r];
            const int L = a.L;
            const size_t nv = (size_t)a.I * a.M;
            // the gather writes a kept input straight into the next history slot
            const bool staged = e < a.start_error || s.n_anderson >= 0;
            const int slot = (s.ring_start + s.ring_n) % L;
            double* am_w = a.am_w + r * a.am_step;
            double* am_d = a.am_d + r * a.am_step;
            double* in_w = staged ? k.w_hist + slot * nv : am_w;
            double* in_d = staged ? k.d_hist + slot * nv : am_d;
            double* w_im = a.contiguous ? a.w_im + r * a.w_step : am_w;
            for (int c = 0; c < a.I; c++) {
                g.w_out[c] = staged || !a.contiguous ? in_w + (size_t)c * a.M : nullptr;
                g.d_out[c] = in_d + (size_t)c * a.M;
            }
            g.scatter = a.contiguous ? 0 : 1;
            it.w = staged ? in_w : w_im;
            it.d = in_d;
            it.out = w_im;
            it.dots = k.dots;
            it.coef = k.coef;
            it.dots_out = nullptr;
            // AndersonMixingDevice::plan (CpuAndersonMixing.cpp:86-160)
            if (e < a.start_error || s.n_anderson >= 0) s.n_anderson = s.n_anderson + 1;
            if (s.n_anderson >= 0) {
                s.n_anderson = min(a.max_hist, s.n_anderson);
                if (s.ring_n == L) s.ring_start = (s.ring_start + 1) % L;
                s.ring_n = min(s.ring_n + 1, L);
                it.na = s.n_anderson;
                for (int n = 0; n <= it.na; n++) {
                    const int row = (s.ring_start + s.ring_n - n - 1 + L) % L;
                    it.wh[n] = k.w_hist + row * nv;
                    it.dh[n] = k.d_hist + row * nv;
                }
                it.dots_out = k.dots + (size_t)((s.dots_start + s.dots_n) % L) * L;
                if (s.dots_n == L) s.dots_start = (s.dots_start + 1) % L;
                s.dots_n = min(s.dots_n + 1, L);
            }
            if (s.n_anderson <= 0) {
                if (old_err < e) {
                    const double m7 = s.mix * 0.7;
                    s.mix = m7 < a.mix_min ? a.mix_min : m7;
                } else {
                    s.mix = s.mix * 1.01;
                }
                it.mode = 1;
                it.mix = s.mix;
            } else {
                it.mode = 2;
                for (int n = 0; n <= s.n_anderson; n++) it.drows[n] = (s.dots_start + s.dots_n - n - 1 + L) % L;
            }
        }
    }
    gather[r] = g;
    st[r] = s;
    status[r] = s.err;
    status[a.NR + r] = (double)s.done;
}

struct Scatter {
    int n;
    const double* src[FTS_MAX_TYPES];
    double* dst[FTS_MAX_TYPES];
};

__global__ void scatter_kernel(Scatter s, int M) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x)
        for (int c = 0; c < s.n; c++) s.dst[c][i] = s.src[c][i];
}

// Leimkuhler-Matthews step of the real fields (lfts.py:705-709):
// w[c] += -lambda[c] * dt * s_c + 0.5 * (prev[c] + cur[c]) * sqrt(s_c)
// blockIdx.y = replica q: w at + q w_step, lambda, prev and cur at + q r_step
struct LmArgs {
    int n;
    double* w[FTS_MAX_TYPES];
    const double* lambda[FTS_MAX_TYPES];
    const double* prev[FTS_MAX_TYPES];
    const double* cur[FTS_MAX_TYPES];
    double dt_s[FTS_MAX_TYPES];
    double sqrt_s[FTS_MAX_TYPES];
    size_t w_step, r_step;
};

__global__ void lm_kernel(LmArgs a, int M) {
    const size_t qw = blockIdx.y * a.w_step, qr = blockIdx.y * a.r_step;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x)
        for (int c = 0; c < a.n; c++)
            a.w[c][qw + i] += -a.lambda[c][qr + i] * a.dt_s[c] +
                              0.5 * (a.prev[c][qr + i] + a.cur[c][qr + i]) * a.sqrt_s[c];
}

int grid_blocks(int M) { return std::min(2048, ceil_div(M, 256)); }

__global__ void add_difference_kernel(double* x, const double* a, const double* b, int M) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) x[i] += a[i] - b[i];
}

// inputs of the residual guess: [w-, first residual, w+ it was evaluated at] (3 M
// values), and the target w+ - that w+ when target is given
__global__ void residual_inputs_kernel(const double* w_real, const double* res, const double* w_imag,
                                       const double* w_start, double* in, double* target, int M) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        in[i] = w_real[i];
        in[M + i] = res[i];
        in[2 * (size_t)M + i] = w_start[i];
        if (target) target[i] = w_imag[i] - w_start[i];
    }
}

__global__ void add_in_place_kernel(double* x, const double* a, int M) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) x[i] += a[i];
}

}  // namespace

SaddlePoint::SaddlePoint(fts_context* ctx, SolverBase* solver, std::vector<AndersonMixingDevice*> ams, int M,
                         const fts_field_theory& ft, int n_blocks, int n_solvents)
    : ctx_(ctx), solver_(solver), ams_(std::move(ams)), M_(M), NR_((int)ams_.size()), n_blocks_(n_blocks),
      n_solvents_(n_solvents) {
    S_ = ft.n_types;
    R_ = ft.n_real;
    I_ = ft.n_imag;
    P_ = ft.n_polymers;
    n_solver_w_ = ft.n_solver_fields;
    FTS_REQUIRE(NR_ >= 1, "at least one replica is required");
    for (auto* a : ams_) FTS_REQUIRE(a != nullptr, "Anderson mixing is required for every replica");
    FTS_REQUIRE(M > 0, "n_grid must be positive");
    FTS_REQUIRE(S_ >= 1 && S_ <= FTS_MAX_TYPES, "n_types must be in [1, 8]");
    FTS_REQUIRE(R_ >= 0 && I_ >= 1 && R_ + I_ <= S_, "invalid numbers of real and imaginary fields");
    FTS_REQUIRE(n_solver_w_ >= 1 && n_solver_w_ <= LC_MAX_OUT, "n_solver_fields must be in [1, 8]");
    FTS_REQUIRE(P_ >= 1, "n_polymers must be positive");
    FTS_REQUIRE(NR_ == 1 || (n_blocks > 0 && n_solvents >= 0), "replicas need the per-replica block count");
    for (int c = 0; c < R_; c++) {
        FTS_REQUIRE(ft.real_idx[c] >= 0 && ft.real_idx[c] < S_, "real_idx out of range");
        real_idx_.push_back(ft.real_idx[c]);
    }
    for (int c = 0; c < I_; c++) {
        FTS_REQUIRE(ft.imag_idx[c] >= 0 && ft.imag_idx[c] < S_, "imag_idx out of range");
        imag_idx_.push_back(ft.imag_idx[c]);
    }
    w_map_.assign(ft.w_map, ft.w_map + (size_t)n_solver_w_ * S_);
    term_off_.assign(ft.phi_term_off, ft.phi_term_off + S_ + 1);
    const int nt = term_off_[S_];
    FTS_REQUIRE(term_off_[0] == 0 && nt >= 0, "invalid phi_term_off");
    for (int i = 0; i < S_; i++) FTS_REQUIRE(term_off_[i + 1] >= term_off_[i], "phi_term_off must be ascending");
    term_kind_.assign(ft.phi_term_kind, ft.phi_term_kind + nt);
    term_index_.assign(ft.phi_term_index, ft.phi_term_index + nt);
    term_scale_.assign(ft.phi_term_scale, ft.phi_term_scale + nt);
    lam_cw_.assign(ft.real_cw, ft.real_cw + (size_t)R_ * S_);
    lam_cp_.assign(ft.real_cp, ft.real_cp + (size_t)R_ * S_);
    lam_c0_.assign(ft.real_c0, ft.real_c0 + R_);
    h_cw_.assign(ft.imag_cw, ft.imag_cw + (size_t)I_ * S_);
    h_cp_.assign(ft.imag_cp, ft.imag_cp + (size_t)I_ * S_);
    h_c0_.assign(ft.imag_c0, ft.imag_c0 + I_);
    dt_scaling_.assign(ft.dt_scaling, ft.dt_scaling + S_);
    h_lin_.assign(ft.h_lin, ft.h_lin + S_);
    h_sq_.assign(ft.h_sq, ft.h_sq + S_);
    pol_weight_.assign(ft.polymer_weight, ft.polymer_weight + P_);
    h_const_ = ft.h_const;
    for (int i : real_idx_) FTS_REQUIRE(dt_scaling_[i] > 0.0, "dt_scaling of the real fields must be positive");
    for (int i : imag_idx_) FTS_REQUIRE(dt_scaling_[i] > 0.0, "dt_scaling of the imaginary fields must be positive");

    const size_t f = sizeof(double) * M * NR_;
    w_aux_.allocate(f * S_);
    phi_.allocate(f * S_);
    w_solver_.allocate(f * n_solver_w_);
    h_.allocate(f * I_);
    lam_.allocate(f * std::max(R_, 1));
    am_w_.allocate(f * I_);
    am_d_.allocate(f * I_);
    noise_prev_.allocate(f * std::max(R_, 1));
    noise_cur_.allocate(f * std::max(R_, 1));
    w_backup_.allocate(f * S_);
    phi_backup_.allocate(f * S_);
    partials_.allocate(sizeof(double) * std::max(2 * FTS_MAX_TYPES, NR_ * S_) * STAT_BLOCKS);
    stats_.allocate(sizeof(double) * 4 * FTS_MAX_TYPES * NR_);
    if (NR_ > 1) {
        gather_host_.allocate_host_pinned(sizeof(GatherRep) * NR_);
        gather_dev_.allocate(sizeof(GatherRep) * NR_);
    }
    FTS_HIP_CHECK(hipMemsetAsync(w_aux_.get(), 0, f * S_, ctx_->stream));
    FTS_HIP_CHECK(hipMemsetAsync(phi_.get(), 0, f * S_, ctx_->stream));
    FTS_HIP_CHECK(hipMemsetAsync(noise_prev_.get(), 0, f * std::max(R_, 1), ctx_->stream));
    FTS_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
}

void SaddlePoint::set_fields(const double* w, bool on_device) {
    field_version_ = ++version_counter_;
    FTS_HIP_CHECK(hipMemcpyAsync(w_aux_.get(), w, sizeof(double) * M_ * S_ * NR_,
                                 on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx_->stream));
    FTS_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
}

void SaddlePoint::get_fields(double* w, bool on_device) {
    FTS_HIP_CHECK(hipMemcpyAsync(w, w_aux_.get(), sizeof(double) * M_ * S_ * NR_,
                                 on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx_->stream));
    FTS_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
}

void SaddlePoint::get_concentration(int type, double* out, bool on_device) {
    FTS_REQUIRE(type >= 0 && type < S_ * NR_, "type index out of range");
    FTS_HIP_CHECK(hipMemcpyAsync(out, phi_.as<double>() + (size_t)type * M_, sizeof(double) * M_,
                                 on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx_->stream));
    FTS_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
}

void SaddlePoint::set_noise(const double* noise, bool on_device) {
    if (R_ == 0) return;
    FTS_HIP_CHECK(hipMemcpyAsync(noise_prev_.get(), noise, sizeof(double) * M_ * R_ * NR_,
                                 on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx_->stream));
    FTS_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
}

void SaddlePoint::get_noise(double* noise, bool on_device) {
    if (R_ == 0) return;
    FTS_HIP_CHECK(hipMemcpyAsync(noise, noise_prev_.get(), sizeof(double) * M_ * R_ * NR_,
                                 on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx_->stream));
    FTS_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
}

void SaddlePoint::backup() {
    backup_field_version_ = field_version_;
    has_backup_ = true;
    const size_t b = sizeof(double) * M_ * S_ * NR_;
    FTS_HIP_CHECK(hipMemcpyAsync(w_backup_.get(), w_aux_.get(), b, hipMemcpyDeviceToDevice, ctx_->stream));
    FTS_HIP_CHECK(hipMemcpyAsync(phi_backup_.get(), phi_.get(), b, hipMemcpyDeviceToDevice, ctx_->stream));
}

void SaddlePoint::restore(const int* replicas) {
    // a partial restore leaves a mix of versions: give it a new one
    field_version_ = replicas ? ++version_counter_ : backup_field_version_;
    const size_t b = sizeof(double) * M_ * S_;
    for (int r = 0; r < NR_; r++) {
        if (replicas && !replicas[r]) continue;
        FTS_HIP_CHECK(hipMemcpyAsync(wa(r), w_backup_.as<double>() + (size_t)r * S_ * M_, b, hipMemcpyDeviceToDevice,
                                     ctx_->stream));
        FTS_HIP_CHECK(hipMemcpyAsync(ph(r), phi_backup_.as<double>() + (size_t)r * S_ * M_, b,
                                     hipMemcpyDeviceToDevice, ctx_->stream));
    }
    FTS_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
}

void SaddlePoint::run_lincomb(int rep, const std::vector<double>& cw, const std::vector<double>& cp,
                              const std::vector<double>& c0, int n_out, double* out, int n_rep, size_t out_step) {
    // inputs: w_aux rows 0..S-1 of the replica, then its phi rows (cp empty: w_aux only);
    // n_rep replicas from rep on, the outputs of replica rep + q at out + q out_step
    LinComb a{};
    a.in_step = (size_t)S_ * M_;
    a.out_step = out_step;
    a.n_out = n_out;
    a.n_in = cp.empty() ? S_ : 2 * S_;
    for (int j = 0; j < S_; j++) {
        a.in[j] = wa(rep) + (size_t)j * M_;
        if (!cp.empty()) a.in[S_ + j] = ph(rep) + (size_t)j * M_;
    }
    for (int r = 0; r < n_out; r++) {
        a.out[r] = out + (size_t)r * M_;
        a.c0[r] = c0.empty() ? 0.0 : c0[r];
        for (int j = 0; j < S_; j++) {
            a.c[r][j] = cw[(size_t)r * S_ + j];
            if (!cp.empty()) a.c[r][S_ + j] = cp[(size_t)r * S_ + j];
        }
    }
    hipLaunchKernelGGL(lincomb_kernel, dim3(grid_blocks(M_), n_rep), dim3(256), 0, ctx_->stream, a, M_);
    FTS_HIP_CHECK(hipGetLastError());
}

void SaddlePoint::row_sums_strided(const double* base, int n, size_t stride, const double* centre, bool square,
                                   double* out_dev) {
    // fixed-order sums of rows base + k stride, k < n (the sums of row_sums)
    const int chunk = std::max(2 * FTS_MAX_TYPES, NR_ * S_);
    for (int r0 = 0; r0 < n; r0 += chunk) {
        const int k = std::min(chunk, n - r0);
        hipLaunchKernelGGL(row_sum_strided_stage1, dim3(STAT_BLOCKS, k), dim3(STAT_THREADS), 0, ctx_->stream,
                           base + (size_t)r0 * stride, stride, M_, centre ? centre + r0 : nullptr, square ? 1 : 0,
                           partials_.as<double>());
        hipLaunchKernelGGL(row_sum_stage2, dim3(k), dim3(STAT_BLOCKS), 0, ctx_->stream, partials_.as<double>(),
                           out_dev + r0);
        FTS_HIP_CHECK(hipGetLastError());
    }
}

void SaddlePoint::row_sums(const std::vector<const double*>& rows, const double* centre, bool square,
                           double* out_dev) {
    // fixed-order sums of any number of rows, 16 rows per launch
    const int n = (int)rows.size();
    for (int r0 = 0; r0 < n; r0 += 2 * FTS_MAX_TYPES) {
        Rows r{};
        const int k = std::min(2 * FTS_MAX_TYPES, n - r0);
        for (int i = 0; i < k; i++) r.row[i] = rows[r0 + i];
        hipLaunchKernelGGL(row_sum_stage1, dim3(STAT_BLOCKS, k), dim3(STAT_THREADS), 0, ctx_->stream, r, M_,
                           centre ? centre + r0 : nullptr, square ? 1 : 0, partials_.as<double>());
        hipLaunchKernelGGL(row_sum_stage2, dim3(k), dim3(STAT_BLOCKS), 0, ctx_->stream, partials_.as<double>(),
                           out_dev + r0);
        FTS_HIP_CHECK(hipGetLastError());
    }
}

void SaddlePoint::concentrations(SolverBase* solver) {
    // monomer fields of the solver (lfts.py:534-545), replica r's at r n_solver_w, all
    // replicas in one launch
    run_lincomb(0, w_map_, {}, {}, n_solver_w_, w_solver_.as<double>(), NR_, (size_t)n_solver_w_ * M_);
    std::vector<const double*> w(n_solver_w_ * NR_);
    for (size_t m = 0; m < w.size(); m++) w[m] = w_solver_.as<double>() + m * M_;
    solver->compute_propagators(w.data(), nullptr, true);
    solver->compute_concentrations();
    // per monomer type, the sums of every replica in one set of launches
    for (int i = 0; i < S_; i++) {
        const int o = term_off_[i], n = term_off_[i + 1] - o;
        solver->sum_concentrations_replicas(n, term_kind_.data() + o, term_index_.data() + o, term_scale_.data() + o,
                                            ph(0) + (size_t)i * M_, NR_, n_blocks_, n_solvents_, (size_t)S_ * M_);
    }
}

void SaddlePoint::field_moments(double* mean, double* mean_sq) {
    std::vector<const double*> rows;
    for (int r = 0; r < NR_; r++)
        for (int j = 0; j < S_; j++) rows.push_back(wa(r) + (size_t)j * M_);
    const int n = (int)rows.size();
    double* st = stats_.as<double>();
    row_sums(rows, nullptr, false, st);
    row_sums(rows, nullptr, true, st + n);
    std::vector<double> h(2 * n);
    FTS_HIP_CHECK(hipMemcpyAsync(h.data(), st, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, ctx_->stream));
    FTS_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
    for (int k = 0; k < n; k++) {
        mean[k] = h[k] / M_;
        mean_sq[k] = h[n + k] / M_;
    }
}

// lfts.py:197-216: H = sum_j (h_sq_j <w_j^2> + h_lin_j <w_j>) - sum_p phi_p/alpha_p ln Q_p + h_const
double SaddlePoint::hamiltonian(int rep, SolverBase* solver) {
    std::vector<const double*> rows;
    for (int j = 0; j < S_; j++) rows.push_back(wa(rep) + (size_t)j * M_);
    double* st = stats_.as<double>();
    row_sums(rows, nullptr, false, st);
    row_sums(rows, nullptr, true, st + S_);
    std::vector<double> m(2 * S_);
    FTS_HIP_CHECK(hipMemcpyAsync(m.data(), st, sizeof(double) * 2 * S_, hipMemcpyDeviceToHost, ctx_->stream));
    FTS_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
    double h = 0.0;
    for (int j = 0; j < S_; j++) h += h_lin_[j] * (m[j] / M_);
    for (int j = 0; j < S_; j++) h += h_sq_[j] * (m[S_ + j] / M_);
    double hp = 0.0;
    for (int p = 0; p < P_; p++) hp -= pol_weight_[p] * std::log(solver->partition(p + rep * P_));
    return hp + h + h_const_;
}

// hamiltonian(rep, solver) for several replicas, with one set of row sums over every
// replica's w_aux rows (the same per-row sums) and one transfer
void SaddlePoint::hamiltonians(const std::vector<int>& reps, SolverBase* solver, std::vector<double>& ham) {
    const int n = NR_ * S_;
    double* st = stats_.as<double>();
    row_sums_strided(w_aux_.as<double>(), n, M_, nullptr, false, st);
    row_sums_strided(w_aux_.as<double>(), n, M_, nullptr, true, st + n);
    std::vector<double> m(2 * n);
    FTS_HIP_CHECK(hipMemcpyAsync(m.data(), st, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, ctx_->stream));
    FTS_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
    for (int rep : reps) {
        const double* sum = m.data() + (size_t)rep * S_;
        const double* sq = m.data() + n + (size_t)rep * S_;
        double h = 0.0;
        for (int j = 0; j < S_; j++) h += h_lin_[j] * (sum[j] / M_);
        for (int j = 0; j < S_; j++) h += h_sq_[j] * (sq[j] / M_);
        double hp = 0.0;
        for (int p = 0; p < P_; p++) hp -= pol_weight_[p] * std::log(solver->partition(p + rep * P_));
        ham[rep] = hp + h + h_const_;
    }
}

// One Anderson step (lfts.py:854-860) of every replica with mix[r], in batched launches:
// the gather of the inputs (into the history slots the mixings keep), the mixings
// (AndersonMixingDevice::calculate_batch) and, for non-contiguous imaginary rows, the
// copy back into w_aux. Each replica's results equal its own step of the unbatched loop.
void SaddlePoint::mix_replicas(const std::vector<char>& mix, const std::vector<double>& old_err,
                               const std::vector<double>& err, bool contiguous) {
    hipStream_t st = ctx_->stream;
    // the previous table copy finished before this iteration's errors were read back
    GatherRep* tab = static_cast<GatherRep*>(gather_host_.host());
    std::vector<double*> out(NR_, nullptr);
    std::vector<const double*> w(NR_, nullptr), d(NR_, nullptr);
    std::vector<char> staged(NR_, 0);
    for (int r = 0; r < NR_; r++) {
        tab[r] = GatherRep{};
        if (!mix[r]) continue;
        double* am_w = am_w_.as<double>() + (size_t)r * I_ * M_;
        double* am_d = am_d_.as<double>() + (size_t)r * I_ * M_;
        staged[r] = ams_[r]->will_insert(err[r]) ? 1 : 0;
        double* in_w = staged[r] ? ams_[r]->staged_w() : am_w;
        double* in_d = staged[r] ? ams_[r]->staged_d() : am_d;
        double* w_im = contiguous ? wa(r) + (size_t)imag_idx_[0] * M_ : am_w;
        for (int c = 0; c < I_; c++) {
            tab[r].w_out[c] = staged[r] || !contiguous ? in_w + (size_t)c * M_ : nullptr;
            tab[r].d_out[c] = in_d + (size_t)c * M_;
        }
        tab[r].scatter = contiguous ? 0 : 1;
        w[r] = staged[r] ? in_w : w_im;
        d[r] = in_d;
        out[r] = w_im;
    }
    FTS_HIP_CHECK(hipMemcpyAsync(gather_dev_.get(), tab, sizeof(GatherRep) * NR_, hipMemcpyHostToDevice, st));
    GatherShared g{};
    g.n = I_;
    for (int c = 0; c < I_; c++) {
        g.wim[c] = wa(0) + (size_t)imag_idx_[c] * M_;
        g.h[c] = h_.as<const double>() + (size_t)c * M_;
        g.scale[c] = dt_scaling_[imag_idx_[c]];
    }
    g.wim_step = (size_t)S_ * M_;
    g.h_step = (size_t)I_ * M_;
    g.am_w = am_w_.as<const double>();
    g.am_step = (size_t)I_ * M_;
    hipLaunchKernelGGL(gather_batch_kernel, dim3(grid_blocks(M_), NR_), dim3(256), 0, st, g,
                       gather_dev_.as<const GatherRep>(), M_);
    FTS_HIP_CHECK(hipGetLastError());
    AndersonMixingDevice::calculate_batch(ctx_, ams_, mix, out, w, d, old_err, err, staged, am_ws_);
    if (!contiguous) {
        hipLaunchKernelGGL(scatter_batch_kernel, dim3(grid_blocks(M_), NR_), dim3(256), 0, st, g,
                           gather_dev_.as<const GatherRep>(), M_);
        FTS_HIP_CHECK(hipGetLastError());
    }
}

SaddlePoint::~SaddlePoint() {
    for (auto& g : rep_a_)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (rep_b_) (void)hipGraphExecDestroy(rep_b_);
    for (auto& g : rep_lm_)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
}

template <typename F>
void SaddlePoint::capture(hipGraphExec_t& exec, F&& issue) {
    hipStream_t st = ctx_->stream;
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
}

// graphs hold no sampled launch events (kernel timing on: the launches run one by one)
bool SaddlePoint::use_replica_graphs() const {
    if (NR_ < 2 || !rep_graphs_ || ctx_->timer.active()) return false;
    for (auto* a : ams_)
        if (a->reduce_memory() || a->max_hist() >= AM_MAX_HIST) return false;
    return true;
}

void SaddlePoint::setup_replica_graphs(bool contiguous) {
    (void)contiguous;
    if (rep_state_.bytes()) return;
    const AndersonMixingDevice& a0 = *ams_[0];
    for (auto* a : ams_)
        FTS_REQUIRE(a->n_var() == a0.n_var() && a->max_hist() == a0.max_hist() &&
                        a->start_error() == a0.start_error() && a->mix_min() == a0.mix_min() &&
                        a->mix_init() == a0.mix_init(),
                    "the replicas' mixings differ in their parameters");
    rep_state_.allocate(sizeof(RepAmState) * NR_);
    rep_const_.allocate(sizeof(RepAmConst) * NR_);
    rep_items_.allocate(sizeof(AmBatchItem) * NR_);
    rep_status_.allocate(sizeof(double) * 2 * NR_);
    rep_status_host_.allocate_host_pinned(sizeof(double) * 2 * NR_);
    std::vector<RepAmConst> k(NR_);
    for (int r = 0; r < NR_; r++) k[r] = RepAmConst{ams_[r]->w_hist(), ams_[r]->d_hist(), ams_[r]->dots_ring(),
                                                    ams_[r]->coef()};
    FTS_HIP_CHECK(hipMemcpyAsync(rep_const_.get(), k.data(), sizeof(RepAmConst) * NR_, hipMemcpyHostToDevice,
                                 ctx_->stream));
    FTS_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
    am_ws_.reserve(NR_);
}

// A: the evaluation, dH/dw of the imaginary fields and their statistics, then the
// device-side plan: errors, convergence and every replica's mixing decisions
void SaddlePoint::issue_graph_a(SolverBase* solver, bool low, double tolerance, bool contiguous) {
    concentrations(solver);
    run_lincomb(0, h_cw_, h_cp_, h_c0_, I_, h_.as<double>(), NR_, (size_t)I_ * M_);
    const int nrows = NR_ * I_;
    double* sd = stats_.as<double>();
    row_sums_strided(h_.as<double>(), nrows, M_, nullptr, false, sd);
    row_sums_strided(h_.as<double>(), nrows, M_, sd, true, sd + nrows);
    const AndersonMixingDevice& a0 = *ams_[0];
    RepPlanArgs a{};
    a.NR = NR_;
    a.I = I_;
    a.M = M_;
    a.L = a0.max_hist() + 1;
    a.max_hist = a0.max_hist();
    a.contiguous = contiguous ? 1 : 0;
    a.low = low ? 1 : 0;
    a.nrows = nrows;
    a.tol = tolerance;
    a.start_error = a0.start_error();
    a.mix_min = a0.mix_min();
    a.mix_init = a0.mix_init();
    a.am_w = am_w_.as<double>();
    a.am_d = am_d_.as<double>();
    a.am_step = (size_t)I_ * M_;
    a.w_im = wa(0) + (size_t)imag_idx_[0] * M_;
    a.w_step = (size_t)S_ * M_;
    hipLaunchKernelGGL(replica_plan_kernel, dim3(ceil_div(NR_, 64)), dim3(64), 0, ctx_->stream, a,
                       stats_.as<const double>(), rep_state_.as<RepAmState>(), rep_const_.as<const RepAmConst>(),
                       rep_items_.as<AmBatchItem>(), gather_dev_.as<GatherRep>(), rep_status_.as<double>());
    FTS_HIP_CHECK(hipGetLastError());
}

// B: the gather into the planned history slots and one batched Anderson step
void SaddlePoint::issue_graph_b(bool contiguous) {
    hipStream_t st = ctx_->stream;
    GatherShared g{};
    g.n = I_;
    for (int c = 0; c < I_; c++) {
        g.wim[c] = wa(0) + (size_t)imag_idx_[c] * M_;
        g.h[c] = h_.as<const double>() + (size_t)c * M_;
        g.scale[c] = dt_scaling_[imag_idx_[c]];
    }
    g.wim_step = (size_t)S_ * M_;
    g.h_step = (size_t)I_ * M_;
    g.am_w = am_w_.as<const double>();
    g.am_step = (size_t)I_ * M_;
    hipLaunchKernelGGL(gather_batch_kernel, dim3(grid_blocks(M_), NR_), dim3(256), 0, st, g,
                       gather_dev_.as<const GatherRep>(), M_);
    FTS_HIP_CHECK(hipGetLastError());
    const int mh = ams_[0]->max_hist();
    AndersonMixingDevice::launch_batch(st, rep_items_.as<const AmBatchItem>(), NR_, ams_[0]->n_var(), mh, mh, mh + 1,
                                       am_ws_.partials.as<double>());
    if (!contiguous) {
        hipLaunchKernelGGL(scatter_batch_kernel, dim3(grid_blocks(M_), NR_), dim3(256), 0, st, g,
                           gather_dev_.as<const GatherRep>(), M_);
        FTS_HIP_CHECK(hipGetLastError());
    }
}

void SaddlePoint::launch_graph_a(SolverBase* solver, bool low, double tolerance, bool contiguous) {
    RepGraph& g = rep_a_[low ? 1 : 0];
    const std::vector<double> sig = solver->launch_signature();
    if (!g.exec || g.solver != solver || g.sig != sig || g.tol != tolerance) {
        // the solver issues its launches into this capture instead of replaying its own graph
        const bool solver_graphs = solver->graphs_on();
        if (solver_graphs) solver->set_graphs(false);
        try {
            capture(g.exec, [&] { issue_graph_a(solver, low, tolerance, contiguous); });
        } catch (...) {
            if (solver_graphs) solver->set_graphs(true);
            throw;
        }
        if (solver_graphs) solver->set_graphs(true);
        g.solver = solver;
        g.sig = sig;
        g.tol = tolerance;
    }
    FTS_HIP_CHECK(hipGraphLaunch(g.exec, ctx_->stream));
    solver->results_changed();
}

void SaddlePoint::launch_graph_b(bool contiguous) {
    if (!rep_b_) capture(rep_b_, [&] { issue_graph_b(contiguous); });
    FTS_HIP_CHECK(hipGraphLaunch(rep_b_, ctx_->stream));
}

// The replica search with every iteration replayed from graphs A and B: the host only
// reads the errors back (one transfer per iteration), switches the precision and stops
// when every replica has converged. The results equal those of the launches issued one
// by one: the plan kernel makes the host loop's decisions in the same arithmetic.
void SaddlePoint::find_replicas_graph(double tolerance, int max_iter, bool contiguous, std::vector<int>& done,
                                      std::vector<double>& err, std::vector<double>& ham) {
    setup_replica_graphs(contiguous);
    hipStream_t st = ctx_->stream;
    hipLaunchKernelGGL(replica_reset_kernel, dim3(ceil_div(NR_, 64)), dim3(64), 0, st, rep_state_.as<RepAmState>(),
                       NR_, ams_[0]->mix_init());
    FTS_HIP_CHECK(hipGetLastError());
    const double* status = static_cast<const double*>(rep_status_host_.host());
    auto read_status = [&] {
        FTS_HIP_CHECK(hipMemcpyAsync(rep_status_host_.host(), rep_status_.get(), sizeof(double) * 2 * NR_,
                                     hipMemcpyDeviceToHost, st));
        FTS_HIP_CHECK(hipStreamSynchronize(st));
        for (int r = 0; r < NR_; r++) {
            err[r] = status[r];
            done[r] = (int)status[NR_ + r];
        }
    };
    bool low = lo_ != nullptr;
    for (int it = 1; it <= max_iter; it++) {
        // the last permitted iteration always evaluates in FP64
        if (it == max_iter) low = false;
        SolverBase* solver = low ? lo_ : solver_;
        launch_graph_a(solver, low, tolerance, contiguous);
        if (it == max_iter) {
            // the Hamiltonian of a replica that stops unconverged is that of the last
            // evaluation, before its last mixing step moves the fields
            read_status();
            std::vector<int> open;
            for (int r = 0; r < NR_; r++)
                if (!done[r]) open.push_back(r);
            if (!open.empty()) hamiltonians(open, solver, ham);
        }
        launch_graph_b(contiguous);
        read_status();
        int active = 0;
        for (int r = 0; r < NR_; r++)
            if (!done[r]) active++;
        if (active == 0) break;
        if (low) {
            bool below = true;
            for (int r = 0; r < NR_; r++)
                if (!done[r] && !(err[r] < lo_switch_)) below = false;
            low = !below;
        }
    }
    // converged replicas kept their fields since their last evaluation, whose partition
    // functions every later evaluation reproduced: their Hamiltonians now
    std::vector<int> conv;
    for (int r = 0; r < NR_; r++)
        if (done[r]) conv.push_back(r);
    if (!conv.empty()) hamiltonians(conv, solver_, ham);
}

void SaddlePoint::find(double tolerance, int max_iter, int verbose, double* log, int* n_iter, double* error,
                       double* hamiltonian_out) {
    FTS_REQUIRE(max_iter >= 1, "max_iter must be positive");
    hipStream_t st = ctx_->stream;
    std::vector<double> err(NR_, 1e20), ham(NR_, std::nan(""));
    std::vector<int> done(NR_, 0);   // iteration at which the replica converged, 0: not yet
    for (auto* a : ams_) a->reset();
    const int row = 3 + P_ + I_;
    const int nrows = NR_ * I_;
    std::vector<double> stats(2 * nrows);
    bool contiguous = true;
    for (int c = 1; c < I_; c++) contiguous = contiguous && imag_idx_[c] == imag_idx_[0] + c;
    // replicas mix in batched launches (mixings with history rings in device memory)
    bool batched = NR_ > 1;
    for (auto* a : ams_) batched = batched && !a->reduce_memory();
    const bool graphs = batched && use_replica_graphs();
    if (graphs) find_replicas_graph(tolerance, max_iter, contiguous, done, err, ham);
    // low-precision phase while the errors are large (none when no low-precision solver is set)
    bool low = lo_ != nullptr;
    for (int it = 1; it <= max_iter && !graphs; it++) {
        // the last permitted iteration always evaluates in FP64: an error measured by
        // the low-precision solver never ends the loop as a converged saddle point
        if (it == max_iter) low = false;
        SolverBase* solver = low ? lo_ : solver_;
        concentrations(solver);
        // dH/dw of the imaginary fields, sign already flipped (lfts.py:219-245), rows
        // h + (r I + c) M of every replica, with their means and deviations
        run_lincomb(0, h_cw_, h_cp_, h_c0_, I_, h_.as<double>(), NR_, (size_t)I_ * M_);
        double* sd = stats_.as<double>();
        row_sums_strided(h_.as<double>(), nrows, M_, nullptr, false, sd);
        row_sums_strided(h_.as<double>(), nrows, M_, sd, true, sd + nrows);
        FTS_HIP_CHECK(hipMemcpyAsync(stats.data(), sd, sizeof(double) * 2 * nrows, hipMemcpyDeviceToHost, st));
        FTS_HIP_CHECK(hipStreamSynchronize(st));
        int active = 0;
        const std::vector<double> prev_err = err;
        std::vector<char> to_mix(NR_, 0);
        std::vector<int> need_h;
        for (int r = 0; r < NR_; r++) {
            if (done[r]) continue;
            const double old_err = err[r];
            double e = 0.0;
            std::vector<double> err_i(I_);
            for (int c = 0; c < I_; c++) {
                err_i[c] = std::sqrt(stats[nrows + r * I_ + c] / M_);
                if (c == 0 || err_i[c] > e || std::isnan(err_i[c])) e = err_i[c];
            }
            err[r] = e;
            // the Hamiltonian and the printed row of lfts.py:834-849, at every iteration
            // (verbose 2) or at the final one
            const bool converged = !low && e < tolerance;
            const bool last = converged || it == max_iter;
            if (NR_ > 1 && last) need_h.push_back(r);
            if (NR_ == 1 && (verbose == 2 || last)) {
                ham[r] = hamiltonian(r, solver);
                if (log && verbose >= 1) {
                    double* lr = log + (size_t)(verbose == 2 ? it - 1 : 0) * row;
                    lr[0] = it;
                    lr[1] = stats[I_ - 1] / M_;   // mass error: mean of the pressure-field derivative
                    lr[2] = ham[r];
                    for (int p = 0; p < P_; p++) lr[3 + p] = solver->partition(p);
                    for (int c = 0; c < I_; c++) lr[3 + P_ + c] = err_i[c];
                }
            }
            if (converged) {
                done[r] = it;
                continue;
            }
            active++;
            bool guess_after_mix = false;
            if (it == 1 && NR_ == 1 && (res_record_ || res_cnn_)) {
                // the residual of the first evaluation: kept for training, and with a
                // trained network its correction replaces this iteration's mixing
                const double* hr0 = h_.as<double>();
                if (res_first_.bytes() == 0) res_first_.allocate(sizeof(double) * M_);
                FTS_HIP_CHECK(hipMemcpyAsync(res_first_.get(), hr0, sizeof(double) * M_, hipMemcpyDeviceToDevice, st));
                has_first_res_ = true;
                if (res_cnn_ && !ams_[r]->will_insert(e)) {
                    apply_residual_guess(*res_cnn_);
                    continue;
                }
                if (res_cnn_) {
                    // the first evaluation also enters the mixing history, so that the
                    // next mixing step has the secant pair (guess - w0, its residual
                    // change); the mixing's own new fields are then replaced by the guess
                    predict_residual_correction(*res_cnn_);
                    if (res_w0_.bytes() == 0) res_w0_.allocate(sizeof(double) * M_);
                    FTS_HIP_CHECK(hipMemcpyAsync(res_w0_.get(), w_aux_.as<double>() + (size_t)imag_idx_[0] * M_,
                                                 sizeof(double) * M_, hipMemcpyDeviceToDevice, st));
                    guess_after_mix = true;
                }
            }
            if (batched) {
                to_mix[r] = 1;
                continue;
            }
            // Anderson mixing of the imaginary fields with d = -dt_scaling * dH/dw (lfts.py:854-860)
            double* am_w = am_w_.as<double>() + (size_t)r * I_ * M_;
            double* am_d = am_d_.as<double>() + (size_t)r * I_ * M_;
            const double* hr = h_.as<double>() + (size_t)r * I_ * M_;
            // when the mixing keeps this iterate, the gather writes it straight into the history slot
            const bool staged = ams_[r]->will_insert(e);
            double* in_w = staged ? ams_[r]->staged_w() : am_w;
            double* in_d = staged ? ams_[r]->staged_d() : am_d;
            Gather g{};
            g.n = I_;
            double* w_im = contiguous ? wa(r) + (size_t)imag_idx_[0] * M_ : am_w;
            for (int c = 0; c < I_; c++) {
                g.src[c] = wa(r) + (size_t)imag_idx_[c] * M_;
                g.h[c] = hr + (size_t)c * M_;
                g.w_out[c] = staged || !contiguous ? in_w + (size_t)c * M_ : nullptr;
                g.d_out[c] = in_d + (size_t)c * M_;
                g.scale[c] = dt_scaling_[imag_idx_[c]];
            }
            hipLaunchKernelGGL(gather_kernel, dim3(grid_blocks(M_)), dim3(256), 0, st, g, M_);
            FTS_HIP_CHECK(hipGetLastError());
            ams_[r]->calculate_new_fields(w_im, staged ? in_w : w_im, in_d, old_err, e, true, staged);
            if (!contiguous) {
                Scatter sc{};
                sc.n = I_;
                for (int c = 0; c < I_; c++) {
                    sc.src[c] = am_w + (size_t)c * M_;
                    sc.dst[c] = wa(r) + (size_t)imag_idx_[c] * M_;
                }
                hipLaunchKernelGGL(scatter_kernel, dim3(grid_blocks(M_)), dim3(256), 0, st, sc, M_);
                FTS_HIP_CHECK(hipGetLastError());
            }
            if (guess_after_mix) {
                double* wi = w_aux_.as<double>() + (size_t)imag_idx_[0] * M_;
                FTS_HIP_CHECK(hipMemcpyAsync(wi, res_w0_.get(), sizeof(double) * M_, hipMemcpyDeviceToDevice, st));
                hipLaunchKernelGGL(add_in_place_kernel, dim3(grid_blocks(M_)), dim3(256), 0, st, wi,
                                   guess_new_.as<const double>(), M_);
                FTS_HIP_CHECK(hipGetLastError());
            }
        }
        // replicas: the Hamiltonians of those that stop here, then one mixing step of the others
        if (!need_h.empty()) hamiltonians(need_h, solver, ham);
        if (batched && active) mix_replicas(to_mix, prev_err, err, contiguous);
        if (active == 0) break;
        if (low) {
            bool below = true;
            for (int r = 0; r < NR_; r++)
                if (!done[r] && !(err[r] < lo_switch_)) below = false;
            low = !below;
        }
    }
    // the residual guess is armed for one search at a time
    res_cnn_ = nullptr;
    res_record_ = false;
    // zero mean of the pressure field (lfts.py:865)
    row_sums_strided(wa(0) + (size_t)(S_ - 1) * M_, NR_, (size_t)S_ * M_, nullptr, false, stats_.as<double>());
    hipLaunchKernelGGL(subtract_mean_kernel, dim3(grid_blocks(M_), NR_), dim3(256), 0, st,
                       wa(0) + (size_t)(S_ - 1) * M_, (size_t)S_ * M_, stats_.as<const double>(), M_);
    FTS_HIP_CHECK(hipGetLastError());
    FTS_HIP_CHECK(hipStreamSynchronize(st));
    for (int r = 0; r < NR_; r++) {
        if (n_iter) n_iter[r] = done[r] ? done[r] : max_iter;
        if (error) error[r] = err[r];
        if (hamiltonian_out) hamiltonian_out[r] = ham[r];
    }
}

void SaddlePoint::langevin_update(const double* noise, bool on_device, double dt) {
    if (R_ == 0) return;
    field_version_ = ++version_counter_;
    hipStream_t st = ctx_->stream;
    FTS_HIP_CHECK(hipMemcpyAsync(noise_cur_.get(), noise, sizeof(double) * M_ * R_ * NR_,
                                 on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    if (use_replica_graphs()) {
        // one graph per order of the two noise buffers, which swap every step
        LmGraph& g = rep_lm_[noise_prev_.get() < noise_cur_.get() ? 0 : 1];
        if (!g.exec || g.prev != noise_prev_.get() || g.dt != dt) {
            capture(g.exec, [&] { issue_langevin(dt); });
            g.prev = noise_prev_.get();
            g.dt = dt;
        }
        FTS_HIP_CHECK(hipGraphLaunch(g.exec, st));
    } else {
        issue_langevin(dt);
    }
    // the current noise becomes the previous one (lfts.py:712)
    std::swap(noise_prev_, noise_cur_);
    FTS_HIP_CHECK(hipStreamSynchronize(st));
}

void SaddlePoint::issue_langevin(double dt) {
    hipStream_t st = ctx_->stream;
    // dH/dw of the real fields (no sign flip) from the current w_aux and phi, then the
    // update, each for every replica in one launch
    run_lincomb(0, lam_cw_, lam_cp_, lam_c0_, R_, lam_.as<double>(), NR_, (size_t)R_ * M_);
    LmArgs a{};
    a.n = R_;
    for (int c = 0; c < R_; c++) {
        const double s = dt_scaling_[real_idx_[c]];
        a.w[c] = wa(0) + (size_t)real_idx_[c] * M_;
        a.lambda[c] = lam_.as<const double>() + (size_t)c * M_;
        a.prev[c] = noise_prev_.as<const double>() + (size_t)c * M_;
        a.cur[c] = noise_cur_.as<const double>() + (size_t)c * M_;
        a.dt_s[c] = dt * s;
        a.sqrt_s[c] = std::sqrt(s);
    }
    a.w_step = (size_t)S_ * M_;
    a.r_step = (size_t)R_ * M_;
    hipLaunchKernelGGL(lm_kernel, dim3(grid_blocks(M_), NR_), dim3(256), 0, st, a, M_);
    FTS_HIP_CHECK(hipGetLastError());
}

void SaddlePoint::prepare_guess(Cnn& cnn) {
    FTS_REQUIRE(NR_ == 1, "the learned guess runs on a single replica");
    FTS_REQUIRE(R_ == 1 && I_ == 1, "the learned guess maps one real field to one imaginary field");
    FTS_REQUIRE(cnn.n_grid() == M_, "the network's grid differs from the simulation grid");
    FTS_REQUIRE(cnn.n_in() == 1, "prepare_guess is for the one-input (delta) guess");
    if (guess_old_.bytes() == 0) guess_old_.allocate(sizeof(double) * M_);
    // f(w_real) is often already there: the previous step's prediction was made from this
    // same field when that step was accepted, and a restored step keeps the prepared one
    const GuessTag tag{&cnn, cnn.version(), field_version_};
    if (!(old_tag_ == tag)) {
        if (new_tag_ == tag) {
            std::swap(guess_old_, guess_new_);
            std::swap(old_tag_, new_tag_);
        } else {
            cnn.predict(w_aux_.as<double>() + (size_t)real_idx_[0] * M_, guess_old_.as<double>(), true);
            old_tag_ = tag;
        }
    }
    prepared_ = true;
}

void SaddlePoint::residual_inputs(double* in, double* target, bool from_backup) {
    FTS_REQUIRE(has_first_res_, "the residual guess needs the first residual of a saddle-point search");
    const double* w = w_aux_.as<const double>();
    const size_t re = (size_t)real_idx_[0] * M_, im = (size_t)imag_idx_[0] * M_;
    // the w+ the first residual was evaluated at: the current one during the search,
    // the backup's (the step's starting point) once the search has moved it
    FTS_REQUIRE(!from_backup || has_backup_, "training the residual guess needs the step's backup");
    const double* start = from_backup ? w_backup_.as<const double>() + im : w + im;
    hipLaunchKernelGGL(residual_inputs_kernel, dim3(grid_blocks(M_)), dim3(256), 0, ctx_->stream, w + re,
                       res_first_.as<const double>(), w + im, start, in, target, M_);
    FTS_HIP_CHECK(hipGetLastError());
}

void SaddlePoint::set_residual_guess(Cnn* cnn, bool record) {
    FTS_REQUIRE(NR_ == 1, "the learned guess runs on a single replica");
    FTS_REQUIRE(R_ == 1 && I_ == 1, "the learned guess maps one real field to one imaginary field");
    FTS_REQUIRE(!cnn || (cnn->n_in() == 3 && cnn->n_grid() == M_),
                "the residual guess needs a three-input network on the simulation grid");
    res_cnn_ = cnn;
    res_record_ = record;
}

void SaddlePoint::predict_residual_correction(Cnn& cnn) {
    if (guess_in_.bytes() == 0) guess_in_.allocate(sizeof(double) * 3 * M_);
    if (guess_new_.bytes() == 0) guess_new_.allocate(sizeof(double) * M_);
    residual_inputs(guess_in_.as<double>(), nullptr, false);
    cnn.predict(guess_in_.as<const double>(), guess_new_.as<double>(), true);
}

void SaddlePoint::apply_residual_guess(Cnn& cnn) {
    predict_residual_correction(cnn);
    hipLaunchKernelGGL(add_in_place_kernel, dim3(grid_blocks(M_)), dim3(256), 0, ctx_->stream,
                       w_aux_.as<double>() + (size_t)imag_idx_[0] * M_, guess_new_.as<const double>(), M_);
    FTS_HIP_CHECK(hipGetLastError());
}

void SaddlePoint::guess_sample(double* out) {
    FTS_REQUIRE(NR_ == 1 && R_ == 1 && I_ == 1, "the residual guess runs on one real and one imaginary field");
    DeviceBuffer tmp(sizeof(double) * 4 * (size_t)M_);
    residual_inputs(tmp.as<double>(), tmp.as<double>() + 3 * (size_t)M_, true);
    FTS_HIP_CHECK(hipMemcpyAsync(out, tmp.get(), sizeof(double) * 4 * (size_t)M_, hipMemcpyDeviceToHost,
                                 ctx_->stream));
    FTS_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
}

void SaddlePoint::predict_guess(Cnn& cnn, int mode) {
    FTS_REQUIRE(NR_ == 1, "the learned guess runs on a single replica");
    FTS_REQUIRE(R_ == 1 && I_ == 1, "the learned guess maps one real field to one imaginary field");
    FTS_REQUIRE(cnn.n_grid() == M_, "the network's grid differs from the simulation grid");
    double* w_real = w_aux_.as<double>() + (size_t)real_idx_[0] * M_;
    double* w_imag = w_aux_.as<double>() + (size_t)imag_idx_[0] * M_;
    FTS_REQUIRE(cnn.n_in() == 1, "the absolute and delta guesses need a one-input network");
    if (mode == 0) {
        cnn.predict(w_real, w_imag, true);
        return;
    }
    FTS_REQUIRE(prepared_, "delta guess needs prepare_guess before the Langevin update");
    if (guess_new_.bytes() == 0) guess_new_.allocate(sizeof(double) * M_);
    cnn.predict(w_real, guess_new_.as<double>(), true);
    new_tag_ = GuessTag{&cnn, cnn.version(), field_version_};
    hipLaunchKernelGGL(add_difference_kernel, dim3(grid_blocks(M_)), dim3(256), 0, ctx_->stream, w_imag,
                       guess_new_.as<const double>(), guess_old_.as<const double>(), M_);
    FTS_HIP_CHECK(hipGetLastError());
    prepared_ = false;
}

double SaddlePoint::train_guess(Cnn& cnn, double lr, int iters, int replay) {
    FTS_REQUIRE(NR_ == 1, "the learned guess runs on a single replica");
    FTS_REQUIRE(R_ == 1 && I_ == 1, "the learned guess maps one real field to one imaginary field");
    FTS_REQUIRE(cnn.n_grid() == M_, "the network's grid differs from the simulation grid");
    FTS_REQUIRE(cnn.n_in() == 1 || cnn.n_in() == 3, "the guess trains one- or three-input networks");
    FTS_REQUIRE(iters >= 1 && replay >= 1, "iterations and replay size must be positive");
    hipStream_t st = ctx_->stream;
    // a sample: the network's inputs, then its target
    const int per = cnn.n_in() + 1;
    if (replay != replay_cap_ || per != replay_per_) {
        replay_.allocate(sizeof(double) * per * (size_t)M_ * replay);
        replay_cap_ = replay;
        replay_per_ = per;
        replay_n_ = replay_head_ = 0;
    }
    // newest sample into the ring
    double* slot = replay_.as<double>() + (size_t)replay_head_ * per * M_;
    if (per == 4) {
        // inputs [w-, first residual, w+ of the step's start], target: the converged w+ minus that start
        residual_inputs(slot, slot + 3 * (size_t)M_, true);
    } else {
        FTS_HIP_CHECK(hipMemcpyAsync(slot, w_aux_.as<double>() + (size_t)real_idx_[0] * M_, sizeof(double) * M_,
                                     hipMemcpyDeviceToDevice, st));
        FTS_HIP_CHECK(hipMemcpyAsync(slot + M_, w_aux_.as<double>() + (size_t)imag_idx_[0] * M_,
                                     sizeof(double) * M_, hipMemcpyDeviceToDevice, st));
    }
    const int newest = replay_head_;
    replay_head_ = (replay_head_ + 1) % replay_cap_;
    replay_n_ = std::min(replay_n_ + 1, replay_cap_);
    double loss = 0.0;
    for (int i = 0; i < iters; i++) {
        int k = newest;
        if (i > 0) {
            // xorshift64 pick among the stored samples
            replay_rng_ ^= replay_rng_ << 13;
            replay_rng_ ^= replay_rng_ >> 7;
            replay_rng_ ^= replay_rng_ << 17;
            k = (int)(replay_rng_ % (unsigned long long)replay_n_);
        }
        const double* sample = replay_.as<double>() + (size_t)k * per * M_;
        const double l = cnn.train_step(sample, sample + (size_t)(per - 1) * M_, true, lr);
        if (i == 0) loss = l;
    }
    return loss;
}

}  // namespace fts
