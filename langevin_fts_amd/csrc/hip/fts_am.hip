// Anderson mixing with device-resident history (CpuAndersonMixing.cpp:76-161,
// AndersonMixing.cpp:19-47, CircularBuffer.cpp:4-54). The n_var-sized history
// rings, the ring of dot-product rows, the small FP64 Gram solve (the unpivoted
// elimination of AndersonMixing::find_an, in the same operation order) and the
// mixing all stay on the device, so a call enqueues work without waiting on it;
// the host keeps only the ring indices and the mixing decisions.
#include <algorithm>
#include <cmath>

#include "fts_internal.hpp"
#include "fts_reduce.hpp"

namespace fts {

namespace {

struct AmDotBatch {
    int n;
    const double* a;
    const double* b[MAX_BATCH];
};

// a . b over this workgroup's share of a grid of gridDim.x workgroups (first stage)
__device__ __forceinline__ double dot_share(const double* a, const double* b, int n_var, double* ws) {
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_var; i += gridDim.x * blockDim.x) acc += a[i] * b[i];
    return block_sum(acc, ws);
}

__global__ void __launch_bounds__(RED_THREADS) am_dot_stage1(AmDotBatch bt, int n_var, double* partials) {
    __shared__ double ws[16];
    const double acc = dot_share(bt.a, bt.b[blockIdx.y], n_var, ws);
    if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = acc;
}

__global__ void __launch_bounds__(RED_BLOCKS) am_dot_stage2(const double* partials, double* out) {
    __shared__ double ws[16];
    double v = partials[blockIdx.x * blockDim.x + threadIdx.x];
    v = block_sum(v, ws);
    if (threadIdx.x == 0) out[blockIdx.x] = v;
}

__device__ __forceinline__ double simple_mix_value(const double* w, const double* d, double mix, int i) {
    return w[i] + mix * d[i];
}

__global__ void am_simple_mix(const double* w, const double* d, double mix, int n, double* out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = simple_mix_value(w, d, mix, i);
}

struct AmMixArgs {
    int n_hist;
    const double* w1;
    const double* e1;
    const double* w2[AM_MAX_HIST];
    const double* e2[AM_MAX_HIST];
    const double* a;   // n_hist coefficients from am_solve
};

struct AmSolveArgs {
    int na;                     // n_anderson
    int L;                      // ring length (row stride)
    int rows[AM_MAX_HIST];      // ring row of D(n, .) for n = 0 (newest) .. na
};

// U a = V with V_i = D(0,0) - D(0,i+1), U_ij = D(0,0) - D(0,i+1) - D(0,j+1) + D(min(i+1,j+1), |i-j|),
// by forward elimination without pivoting and back substitution (AndersonMixing.cpp:19-47,
// CpuAndersonMixing.cpp:128-147), with the host code's operation order and no contraction into
// fused multiply-adds, so the coefficients are those the host would compute. Row j of the
// elimination belongs to thread j; one workgroup of AM_MAX_HIST threads.
__device__ __forceinline__ void am_solve_body(const double* ring, int na, int L, const int* rows, double* coef,
                                              double* lds) {
#pragma clang fp contract(off)
    // LDS: D (na+1 rows of L), u (na x na), v, x
    const int t = threadIdx.x;
    double* D = lds;
    double (*u)[AM_MAX_HIST] = reinterpret_cast<double (*)[AM_MAX_HIST]>(D + (size_t)(na + 1) * L);
    double* v = reinterpret_cast<double*>(u + na);
    double* x = v + AM_MAX_HIST;
    // the rows D(n, .), n = 0 (newest) .. na, in one coalesced pass
    for (int e = t; e < (na + 1) * L; e += blockDim.x) {
        const int n = e / L;
        D[e] = ring[(size_t)rows[n] * L + (e - n * L)];
    }
    __syncthreads();
    const double* d0 = D;
    if (t < na) {
        v[t] = d0[0] - d0[t + 1];
        for (int j = 0; j < na; j++) {
            const int m = t < j ? t + 1 : j + 1;
            const int c = t > j ? t - j : j - t;
            u[t][j] = d0[0] - d0[t + 1] - d0[j + 1] + D[m * L + c];
        }
    }
    __syncthreads();
    for (int i = 0; i < na; i++) {
        if (t > i && t < na) {
            const double factor = u[t][i] / u[i][i];
            v[t] = v[t] - v[i] * factor;
            for (int k = i + 1; k < na; k++) u[t][k] = u[t][k] - u[i][k] * factor;
        }
        __syncthreads();
    }
    if (t == 0) {
        x[na - 1] = v[na - 1] / u[na - 1][na - 1];
        for (int i = na - 2; i >= 0; i--) {
            double s = 0.0;
            for (int j = i + 1; j < na; j++) s = s + u[i][j] * x[j];
            x[i] = (v[i] - s) / u[i][i];
        }
        for (int i = 0; i < na; i++) coef[i] = x[i];
    }
}

__global__ void __launch_bounds__(AM_MAX_HIST) am_solve(const double* ring, AmSolveArgs a, double* coef) {
    extern __shared__ double lds[];
    am_solve_body(ring, a.na, a.L, a.rows, coef, lds);
}

// w_new = (w_0 + d_0) + sum_i a_i ((w_{i+1} + d_{i+1}) - (w_0 + d_0)), same operation order
__global__ void am_mix(AmMixArgs args, int n, double* out) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const double w1 = args.w1[j], e1 = args.e1[j];
        double v = w1 + e1;
        for (int i = 0; i < args.n_hist; i++) v += args.a[i] * (args.w2[i][j] + args.e2[i][j] - w1 - e1);
        out[j] = v;
    }
}

// ---- batched steps: blockIdx.z (dot products) or the last grid index is the mixing
__global__ void __launch_bounds__(RED_THREADS) am_dot_batch_stage1(const AmBatchItem* items, int n_var,
                                                                   double* partials) {
    __shared__ double ws[16];
    const AmBatchItem& it = items[blockIdx.z];
    const int n = blockIdx.y;
    if (n > it.na) return;
    const double acc = dot_share(it.d, it.dh[n], n_var, ws);
    if (threadIdx.x == 0) partials[((size_t)blockIdx.z * AM_MAX_HIST + n) * RED_BLOCKS + blockIdx.x] = acc;
}

__global__ void __launch_bounds__(RED_BLOCKS) am_dot_batch_stage2(const AmBatchItem* items, const double* partials) {
    __shared__ double ws[16];
    const AmBatchItem& it = items[blockIdx.y];
    const int n = blockIdx.x;
    if (n > it.na) return;
    double v = partials[((size_t)blockIdx.y * AM_MAX_HIST + n) * RED_BLOCKS + threadIdx.x];
    v = block_sum(v, ws);
    if (threadIdx.x == 0) it.dots_out[n] = v;
}

__global__ void __launch_bounds__(AM_MAX_HIST) am_solve_batch(const AmBatchItem* items) {
    extern __shared__ double lds[];
    const AmBatchItem& it = items[blockIdx.x];
    if (it.mode != 2) return;
    am_solve_body(it.dots, it.na, it.L, it.drows, it.coef, lds);
}

__global__ void am_mix_batch(const AmBatchItem* items, int n) {
    const AmBatchItem& it = items[blockIdx.y];
    if (it.mode == 1) {
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
            it.out[i] = simple_mix_value(it.w, it.d, it.mix, i);
    } else if (it.mode == 2) {
        // the expression of am_mix, history rows n = 0 (newest) .. na
        for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
            const double w1 = it.wh[0][j], e1 = it.dh[0][j];
            double v = w1 + e1;
            for (int i = 0; i < it.na; i++) v += it.coef[i] * (it.wh[i + 1][j] + it.dh[i + 1][j] - w1 - e1);
            it.out[j] = v;
        }
    }
}

size_t solve_lds_bytes(int na, int L) {
    return sizeof(double) * ((size_t)(na + 1) * L + (size_t)na * AM_MAX_HIST + 2 * AM_MAX_HIST);
}

}  // namespace

void AmBatchWorkspace::reserve(int n) {
    if (n <= capacity) return;
    host_table.allocate_host_pinned(sizeof(AmBatchItem) * n);
    table.allocate(sizeof(AmBatchItem) * n);
    partials.allocate(sizeof(double) * n * AM_MAX_HIST * RED_BLOCKS);
    capacity = n;
}

AndersonMixingDevice::AndersonMixingDevice(fts_context* ctx, int n_var, int max_hist, double start_error,
                                           double mix_min, double mix_init, bool reduce_memory)
    : ctx_(ctx),
      n_var_(n_var),
      max_hist_(max_hist),
      n_anderson_(-1),
      start_error_(start_error),
      mix_min_(mix_min),
      mix_(mix_init),
      mix_init_(mix_init) {
    FTS_REQUIRE(n_var > 0, "n_var must be positive");
    FTS_REQUIRE(max_hist >= 1 && max_hist < AM_MAX_HIST, "max_hist must be in [1, 63]");
    const size_t rows = max_hist + 1;
    // reduce-memory form (CudaAndersonMixingReduceMemory.cu): the history rings live in pinned host memory
    if (reduce_memory) {
        w_hist_.allocate_host_pinned(sizeof(double) * rows * n_var);
        d_hist_.allocate_host_pinned(sizeof(double) * rows * n_var);
    } else {
        w_hist_.allocate(sizeof(double) * rows * n_var);
        d_hist_.allocate(sizeof(double) * rows * n_var);
    }
    in_w_.allocate(sizeof(double) * n_var);
    in_d_.allocate(sizeof(double) * n_var);
    out_.allocate(sizeof(double) * n_var);
    partials_.allocate(sizeof(double) * MAX_BATCH * RED_BLOCKS);
    dots_ring_.allocate(sizeof(double) * rows * rows);
    coef_.allocate(sizeof(double) * AM_MAX_HIST);
    FTS_HIP_CHECK(hipMemset(dots_ring_.get(), 0, dots_ring_.bytes()));
    // the solve's LDS grows with the history (up to about 66 KB at 63)
    const size_t solve_lds = sizeof(double) * (rows * rows + (rows - 1) * AM_MAX_HIST + 2 * AM_MAX_HIST);
    if (solve_lds > 64 * 1024) {
        FTS_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&am_solve),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)solve_lds));
        FTS_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&am_solve_batch),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)solve_lds));
    }
    reset();
}

void AndersonMixingDevice::reset() {
    mix_ = mix_init_;
    n_anderson_ = -1;
    ring_start_ = ring_n_ = 0;
    dots_start_ = dots_n_ = 0;
}

double* AndersonMixingDevice::staged_w() const {
    return w_hist_.as<double>() + (size_t)((ring_start_ + ring_n_) % (max_hist_ + 1)) * n_var_;
}

double* AndersonMixingDevice::staged_d() const {
    return d_hist_.as<double>() + (size_t)((ring_start_ + ring_n_) % (max_hist_ + 1)) * n_var_;
}

AmBatchItem AndersonMixingDevice::plan(const double* w, const double* d, double* out, double old_err, double err) {
    const int L = max_hist_ + 1;
    AmBatchItem it{};
    it.mode = 1;
    it.na = -1;
    it.L = L;
    it.w = w;
    it.d = d;
    it.out = out;
    it.dots = dots_ring_.as<const double>();
    it.coef = coef_.as<double>();
    auto ring_row = [&](int n) { return (ring_start_ + ring_n_ - n - 1 + L) % L; };
    auto dots_row = [&](int n) { return (dots_start_ + dots_n_ - n - 1 + L) % L; };

    if (err < start_error_ || n_anderson_ >= 0) n_anderson_ = n_anderson_ + 1;
    if (n_anderson_ >= 0) {
        n_anderson_ = std::min(max_hist_, n_anderson_);
        // the input enters the ring at slot (ring_start + ring_n) % L
        if (ring_n_ == L) ring_start_ = (ring_start_ + 1) % L;
        ring_n_ = std::min(ring_n_ + 1, L);
        it.na = n_anderson_;
        for (int n = 0; n <= it.na; n++) {
            it.wh[n] = w_hist_.as<const double>() + (size_t)ring_row(n) * n_var_;
            it.dh[n] = d_hist_.as<const double>() + (size_t)ring_row(n) * n_var_;
        }
        // dots[n] = d . d_hist[n], n = 0..n_anderson (no volume weights), into the next row
        // of the ring of dot-product rows
        it.dots_out = dots_ring_.as<double>() + (size_t)((dots_start_ + dots_n_) % L) * L;
        if (dots_n_ == L) dots_start_ = (dots_start_ + 1) % L;
        dots_n_ = std::min(dots_n_ + 1, L);
    }
    if (n_anderson_ <= 0) {
        if (old_err < err) mix_ = std::max(mix_ * 0.7, mix_min_);
        else mix_ = mix_ * 1.01;
        it.mode = 1;
        it.mix = mix_;
    } else {
        it.mode = 2;
        for (int n = 0; n <= n_anderson_; n++) it.drows[n] = dots_row(n);
    }
    return it;
}

void AndersonMixingDevice::calculate_new_fields(double* w_new, const double* w, const double* d, double old_err,
                                                double err, bool on_device, bool staged) {
    FTS_REQUIRE(!staged || (on_device && will_insert(err) && w == staged_w() && d == staged_d()),
                "staged fields must sit in the next history slot");
    hipStream_t st = ctx_->stream;
    const size_t nb = sizeof(double) * n_var_;
    const double* dw = w;
    const double* dd = d;
    if (!on_device) {
        FTS_HIP_CHECK(hipMemcpyAsync(in_w_.get(), w, nb, hipMemcpyHostToDevice, st));
        FTS_HIP_CHECK(hipMemcpyAsync(in_d_.get(), d, nb, hipMemcpyHostToDevice, st));
        dw = in_w_.as<double>();
        dd = in_d_.as<double>();
    }
    // the rings may be pinned host memory: the runtime picks the direction
    if (!staged && will_insert(err)) {
        FTS_HIP_CHECK(hipMemcpyAsync(staged_w(), dw, nb, hipMemcpyDefault, st));
        FTS_HIP_CHECK(hipMemcpyAsync(staged_d(), dd, nb, hipMemcpyDefault, st));
    }
    double* out = on_device ? w_new : out_.as<double>();
    const AmBatchItem it = plan(dw, dd, out, old_err, err);
    for (int i0 = 0; i0 <= it.na; i0 += MAX_BATCH) {
        AmDotBatch b{};
        b.n = std::min(MAX_BATCH, it.na + 1 - i0);
        b.a = dd;
        for (int k = 0; k < b.n; k++) b.b[k] = it.dh[i0 + k];
        hipLaunchKernelGGL(am_dot_stage1, dim3(RED_BLOCKS, b.n), dim3(RED_THREADS), 0, st, b, n_var_,
                           partials_.as<double>());
        hipLaunchKernelGGL(am_dot_stage2, dim3(b.n), dim3(RED_BLOCKS), 0, st, partials_.as<double>(),
                           it.dots_out + i0);
        FTS_HIP_CHECK(hipGetLastError());
    }
    const int gb = std::min(1024, ceil_div(n_var_, 256));
    if (it.mode == 1) {
        hipLaunchKernelGGL(am_simple_mix, dim3(gb), dim3(256), 0, st, dw, dd, it.mix, n_var_, out);
        FTS_HIP_CHECK(hipGetLastError());
    } else {
        const int na = it.na;
        AmSolveArgs sa{};
        sa.na = na;
        sa.L = it.L;
        for (int n = 0; n <= na; n++) sa.rows[n] = it.drows[n];
        hipLaunchKernelGGL(am_solve, dim3(1), dim3(AM_MAX_HIST), solve_lds_bytes(na, it.L), st,
                           dots_ring_.as<const double>(), sa, coef_.as<double>());
        FTS_HIP_CHECK(hipGetLastError());
        AmMixArgs args{};
        args.n_hist = na;
        args.w1 = it.wh[0];
        args.e1 = it.dh[0];
        for (int i = 0; i < na; i++) {
            args.w2[i] = it.wh[i + 1];
            args.e2[i] = it.dh[i + 1];
        }
        args.a = coef_.as<const double>();
        hipLaunchKernelGGL(am_mix, dim3(gb), dim3(256), 0, st, args, n_var_, out);
        FTS_HIP_CHECK(hipGetLastError());
    }
    if (!on_device) {
        FTS_HIP_CHECK(hipMemcpyAsync(w_new, out, nb, hipMemcpyDeviceToHost, st));
        FTS_HIP_CHECK(hipStreamSynchronize(st));
    }
}

void AndersonMixingDevice::calculate_batch(fts_context* ctx, const std::vector<AndersonMixingDevice*>& ams,
                                           const std::vector<char>& active, const std::vector<double*>& out,
                                           const std::vector<const double*>& w, const std::vector<const double*>& d,
                                           const std::vector<double>& old_err, const std::vector<double>& err,
                                           const std::vector<char>& staged, AmBatchWorkspace& ws) {
    const int n = (int)ams.size();
    FTS_REQUIRE(n >= 1 && (int)active.size() == n && (int)out.size() == n && (int)w.size() == n &&
                    (int)d.size() == n && (int)old_err.size() == n && (int)err.size() == n &&
                    (int)staged.size() == n,
                "one entry per mixing");
    hipStream_t st = ctx->stream;
    // the pinned table is rewritten only after its previous copy has left it
    if (ws.copied) FTS_HIP_CHECK(hipEventSynchronize(ws.copied));
    else FTS_HIP_CHECK(hipEventCreateWithFlags(&ws.copied, hipEventDisableTiming));
    ws.reserve(n);
    AmBatchItem* items = static_cast<AmBatchItem*>(ws.host_table.host());
    const int n_var = ams[0]->n_var_;
    int max_na = -1, max_solve = -1, L = ams[0]->max_hist_ + 1;
    bool any = false;
    for (int r = 0; r < n; r++) {
        AndersonMixingDevice* a = ams[r];
        FTS_REQUIRE(a->n_var_ == n_var && a->max_hist_ + 1 == L, "batched mixings share n_var and the history");
        if (!active[r]) {
            items[r] = AmBatchItem{};
            items[r].mode = 0;
            items[r].na = -1;
            continue;
        }
        FTS_REQUIRE(!a->reduce_memory() && a->will_insert(err[r]) == (staged[r] != 0) &&
                        (!staged[r] || (w[r] == a->staged_w() && d[r] == a->staged_d())),
                    "a batched input that the mixing keeps must sit in its next history slot");
        items[r] = a->plan(w[r], d[r], out[r], old_err[r], err[r]);
        max_na = std::max(max_na, items[r].na);
        if (items[r].mode == 2) max_solve = std::max(max_solve, items[r].na);
        any = true;
    }
    if (!any) return;
    FTS_HIP_CHECK(hipMemcpyAsync(ws.table.get(), items, sizeof(AmBatchItem) * n, hipMemcpyHostToDevice, st));
    FTS_HIP_CHECK(hipEventRecord(ws.copied, st));
    launch_batch(st, ws.table.as<const AmBatchItem>(), n, n_var, max_na, max_solve, L, ws.partials.as<double>());
}

void AndersonMixingDevice::launch_batch(hipStream_t st, const AmBatchItem* dev, int n, int n_var, int max_na,
                                        int max_solve, int L, double* partials) {
    FTS_REQUIRE(max_na < AM_MAX_HIST && max_solve < AM_MAX_HIST, "history beyond the batched kernels' table");
    if (max_na >= 0) {
        hipLaunchKernelGGL(am_dot_batch_stage1, dim3(RED_BLOCKS, max_na + 1, n), dim3(RED_THREADS), 0, st, dev,
                           n_var, partials);
        hipLaunchKernelGGL(am_dot_batch_stage2, dim3(max_na + 1, n), dim3(RED_BLOCKS), 0, st, dev,
                           (const double*)partials);
        FTS_HIP_CHECK(hipGetLastError());
    }
    if (max_solve > 0) {
        hipLaunchKernelGGL(am_solve_batch, dim3(n), dim3(AM_MAX_HIST), solve_lds_bytes(max_solve, L), st, dev);
        FTS_HIP_CHECK(hipGetLastError());
    }
    const int gb = std::min(1024, ceil_div(n_var, 256));
    hipLaunchKernelGGL(am_mix_batch, dim3(gb, n), dim3(256), 0, st, dev, n_var);
    FTS_HIP_CHECK(hipGetLastError());
}

}  // namespace fts
