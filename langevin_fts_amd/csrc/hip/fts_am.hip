// Anderson mixing with device-resident history (CpuAndersonMixing.cpp:76-161,
// AndersonMixing.cpp:19-47, CircularBuffer.cpp:4-54). The n_var-sized history
// rings and the mixing stay in HBM; only the (n_anderson+1) dot products cross to
// the host, where the small FP64 Gram system is solved by the unpivoted
// elimination of AndersonMixing::find_an.
#include <algorithm>
#include <cmath>

#include "fts_internal.hpp"
#include "fts_reduce.hpp"

namespace fts {

namespace {

constexpr int AM_MAX_HIST = 64;

struct AmDotBatch {
    int n;
    const double* a;
    const double* b[MAX_BATCH];
};

__global__ void __launch_bounds__(RED_THREADS) am_dot_stage1(AmDotBatch bt, int n_var, double* partials) {
    __shared__ double ws[16];
    const double* b = bt.b[blockIdx.y];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_var; i += gridDim.x * blockDim.x) acc += bt.a[i] * b[i];
    acc = block_sum(acc, ws);
    if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = acc;
}

__global__ void __launch_bounds__(RED_BLOCKS) am_dot_stage2(const double* partials, double* out) {
    __shared__ double ws[16];
    double v = partials[blockIdx.x * blockDim.x + threadIdx.x];
    v = block_sum(v, ws);
    if (threadIdx.x == 0) out[blockIdx.x] = v;
}

__global__ void am_simple_mix(const double* w, const double* d, double mix, int n, double* out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = w[i] + mix * d[i];
}

struct AmMixArgs {
    int n_hist;
    const double* w1;
    const double* e1;
    const double* w2[AM_MAX_HIST];
    const double* e2[AM_MAX_HIST];
    double a[AM_MAX_HIST];
};

// w_new = (w_0 + d_0) + sum_i a_i ((w_{i+1} + d_{i+1}) - (w_0 + d_0)), same operation order
__global__ void am_mix(AmMixArgs args, int n, double* out) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const double w1 = args.w1[j], e1 = args.e1[j];
        double v = w1 + e1;
        for (int i = 0; i < args.n_hist; i++) v += args.a[i] * (args.w2[i][j] + args.e2[i][j] - w1 - e1);
        out[j] = v;
    }
}

}  // namespace

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
    dots_d_.allocate(sizeof(double) * rows);
    dots_.assign(rows * rows, 0.0);
    reset();
}

void AndersonMixingDevice::reset() {
    mix_ = mix_init_;
    n_anderson_ = -1;
    ring_start_ = ring_n_ = 0;
    dots_start_ = dots_n_ = 0;
}

void AndersonMixingDevice::calculate_new_fields(double* w_new, const double* w, const double* d, double old_err,
                                                double err, bool on_device) {
    hipStream_t st = ctx_->stream;
    const int L = max_hist_ + 1;
    const size_t nb = sizeof(double) * n_var_;
    const double* dw = w;
    const double* dd = d;
    if (!on_device) {
        FTS_HIP_CHECK(hipMemcpyAsync(in_w_.get(), w, nb, hipMemcpyHostToDevice, st));
        FTS_HIP_CHECK(hipMemcpyAsync(in_d_.get(), d, nb, hipMemcpyHostToDevice, st));
        dw = in_w_.as<double>();
        dd = in_d_.as<double>();
    }
    auto ring_row = [&](int n) { return (ring_start_ + ring_n_ - n - 1 + L) % L; };
    auto dots_row = [&](int n) { return &dots_[(size_t)((dots_start_ + dots_n_ - n - 1 + L) % L) * L]; };

    if (err < start_error_ || n_anderson_ >= 0) n_anderson_ = n_anderson_ + 1;
    if (n_anderson_ >= 0) {
        n_anderson_ = std::min(max_hist_, n_anderson_);
        const int slot = (ring_start_ + ring_n_) % L;
        // the rings may be pinned host memory: the runtime picks the direction
        FTS_HIP_CHECK(hipMemcpyAsync(w_hist_.as<double>() + (size_t)slot * n_var_, dw, nb, hipMemcpyDefault, st));
        FTS_HIP_CHECK(hipMemcpyAsync(d_hist_.as<double>() + (size_t)slot * n_var_, dd, nb, hipMemcpyDefault, st));
        if (ring_n_ == L) ring_start_ = (ring_start_ + 1) % L;
        ring_n_ = std::min(ring_n_ + 1, L);
        // dots[i] = d . d_hist[i], i = 0..n_anderson (no volume weights)
        for (int i0 = 0; i0 <= n_anderson_; i0 += MAX_BATCH) {
            AmDotBatch b{};
            b.n = std::min(MAX_BATCH, n_anderson_ + 1 - i0);
            b.a = dd;
            for (int k = 0; k < b.n; k++) b.b[k] = d_hist_.as<double>() + (size_t)ring_row(i0 + k) * n_var_;
            hipLaunchKernelGGL(am_dot_stage1, dim3(RED_BLOCKS, b.n), dim3(RED_THREADS), 0, st, b, n_var_,
                               partials_.as<double>());
            hipLaunchKernelGGL(am_dot_stage2, dim3(b.n), dim3(RED_BLOCKS), 0, st, partials_.as<double>(),
                               dots_d_.as<double>() + i0);
            FTS_HIP_CHECK(hipGetLastError());
        }
        std::vector<double> row(L, 0.0);
        FTS_HIP_CHECK(hipMemcpyAsync(row.data(), dots_d_.get(), sizeof(double) * (n_anderson_ + 1),
                                     hipMemcpyDeviceToHost, st));
        FTS_HIP_CHECK(hipStreamSynchronize(st));
        const int ds = (dots_start_ + dots_n_) % L;
        std::copy(row.begin(), row.end(), dots_.begin() + (size_t)ds * L);
        if (dots_n_ == L) dots_start_ = (dots_start_ + 1) % L;
        dots_n_ = std::min(dots_n_ + 1, L);
    }
    double* out = on_device ? w_new : out_.as<double>();
    const int gb = std::min(1024, ceil_div(n_var_, 256));
    if (n_anderson_ <= 0) {
        if (old_err < err) mix_ = std::max(mix_ * 0.7, mix_min_);
        else mix_ = mix_ * 1.01;
        hipLaunchKernelGGL(am_simple_mix, dim3(gb), dim3(256), 0, st, dw, dd, mix_, n_var_, out);
        FTS_HIP_CHECK(hipGetLastError());
    } else {
        const int na = n_anderson_;
        std::vector<double> u((size_t)na * na), v(na), a(na);
        const double* d0 = dots_row(0);
        for (int i = 0; i < na; i++) {
            v[i] = d0[0] - d0[i + 1];
            for (int j = 0; j < na; j++)
                u[(size_t)i * na + j] =
                    d0[0] - d0[i + 1] - d0[j + 1] + dots_row(std::min(i + 1, j + 1))[std::abs(i - j)];
        }
        // AndersonMixing::find_an: Gaussian elimination without pivoting
        for (int i = 0; i < na; i++)
            for (int j = i + 1; j < na; j++) {
                const double factor = u[(size_t)j * na + i] / u[(size_t)i * na + i];
                v[j] = v[j] - v[i] * factor;
                for (int k = i + 1; k < na; k++) u[(size_t)j * na + k] = u[(size_t)j * na + k] - u[(size_t)i * na + k] * factor;
            }
        a[na - 1] = v[na - 1] / u[(size_t)(na - 1) * na + na - 1];
        for (int i = na - 2; i >= 0; i--) {
            double t = 0.0;
            for (int j = i + 1; j < na; j++) t = t + u[(size_t)i * na + j] * a[j];
            a[i] = (v[i] - t) / u[(size_t)i * na + i];
        }
        AmMixArgs args{};
        args.n_hist = na;
        args.w1 = w_hist_.as<double>() + (size_t)ring_row(0) * n_var_;
        args.e1 = d_hist_.as<double>() + (size_t)ring_row(0) * n_var_;
        for (int i = 0; i < na; i++) {
            args.w2[i] = w_hist_.as<double>() + (size_t)ring_row(i + 1) * n_var_;
            args.e2[i] = d_hist_.as<double>() + (size_t)ring_row(i + 1) * n_var_;
            args.a[i] = a[i];
        }
        hipLaunchKernelGGL(am_mix, dim3(gb), dim3(256), 0, st, args, n_var_, out);
        FTS_HIP_CHECK(hipGetLastError());
    }
    if (!on_device) {
        FTS_HIP_CHECK(hipMemcpyAsync(w_new, out, nb, hipMemcpyDeviceToHost, st));
        FTS_HIP_CHECK(hipStreamSynchronize(st));
    }
}

}  // namespace fts
