// Reductions and elementwise kernels of the propagator path. Reductions are
// two-stage with a fixed partition (RED_BLOCKS x RED_THREADS first stage, one
// workgroup second stage) and a fixed wave64 shuffle tree, so results are
// bit-reproducible run to run. All accumulation is FP64.
#pragma once

#include "fts_internal.hpp"

namespace fts {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// sum over the workgroup; result valid in thread 0. Uses 16 doubles of LDS.
__device__ __forceinline__ double block_sum(double v, double* ws) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) ws[wid] = v;
    __syncthreads();
    const int nw = (blockDim.x + 63) >> 6;
    double r = 0.0;
    if (wid == 0) {
        r = lane < nw ? ws[lane] : 0.0;
        r = wave_sum(r);
    }
    __syncthreads();
    return r;
}

template <typename T>
__global__ void __launch_bounds__(RED_THREADS) dot_stage1(DotBatch<T> b, int M, double dV, const double* mask,
                                                          double* partials) {
    __shared__ double ws[16];
    const DotItem<T>& it = b.it[blockIdx.y];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        double v = dV;
        if (mask) v *= mask[i];
        if (it.a) v *= (double)it.a[i];
        if (it.b) v *= (double)it.b[i];
        if (it.w) v *= exp((double)it.wcoef * (double)it.w[i]);
        acc += v;
    }
    acc = block_sum(acc, ws);
    if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = acc;
}

template <typename T>
__global__ void __launch_bounds__(RED_BLOCKS) dot_stage2(DotBatch<T> b, const double* partials, double* out) {
    __shared__ double ws[16];
    double v = partials[blockIdx.x * blockDim.x + threadIdx.x];
    v = block_sum(v, ws);
    if (threadIdx.x == 0) out[blockIdx.x] = v * b.it[blockIdx.x].scale;
}

// SimpsonRule.h:10-59
__device__ __forceinline__ double simpson_coeff(int n, int N) {
    if (N == 0) return 1.0;
    if (N == 1) return 0.5;
    if (N == 3) return (n == 0 || n == 3) ? 3.0 / 8.0 : 9.0 / 8.0;
    if (N % 2 == 0) {
        if (n == 0 || n == N) return 1.0 / 3.0;
        return (n % 2 == 1) ? 4.0 / 3.0 : 2.0 / 3.0;
    }
    if (n == 0) return 1.0 / 3.0;
    if (n <= N - 4) return (n % 2 == 1) ? 4.0 / 3.0 : 2.0 / 3.0;
    if (n == N - 3) return 3.0 / 8.0 + 1.0 / 3.0;
    if (n == N - 2 || n == N - 1) return 9.0 / 8.0;
    return 3.0 / 8.0;
}

// Concentration of one computation block (CpuComputationContinuous.cpp:444-465,
// CpuComputationDiscrete.cpp:667-690): continuous sum_n s_n qL[NL-n] qR[n];
// discrete sum_{n>=1} qL[NL-n+1] qR[n] / exp(-w ds); times norm / Q[polymer].
template <typename T>
struct PhiItem {
    const T* qL;
    const T* qR;
    const T* w;     // discrete: field of the block's monomer
    T wcoef;        // discrete: ds (multiplies by exp(+w ds) = 1/E)
    int NL, NR;
    int polymer;
    double norm;    // ds * vf / alpha * n_repeated
    double* out;
};

template <typename T>
struct PhiBatch {
    int n;
    PhiItem<T> it[MAX_BATCH];
};

template <typename T, bool DISCRETE>
__global__ void __launch_bounds__(256) phi_kernel(PhiBatch<T> b, int M, const double* Q) {
    const PhiItem<T>& it = b.it[blockIdx.y];
    const double norm = it.norm / Q[it.polymer];
    const size_t Ms = (size_t)M;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        double acc = 0.0;
        if (it.NR == 0) {  // block piece without segments
            it.out[i] = 0.0;
            continue;
        }
        if constexpr (!DISCRETE) {
            for (int n = 0; n <= it.NR; n++)
                acc += simpson_coeff(n, it.NR) * (double)it.qL[(size_t)(it.NL - n) * Ms + i] *
                       (double)it.qR[(size_t)n * Ms + i];
        } else {
            for (int n = 1; n <= it.NR; n++)
                acc += (double)it.qL[(size_t)(it.NL - n + 1) * Ms + i] * (double)it.qR[(size_t)n * Ms + i];
            acc *= exp((double)it.wcoef * (double)it.w[i]);
        }
        it.out[i] = acc * norm;
    }
}

// Solvent (CpuComputationContinuous.cpp:425-436, CpuComputationDiscrete.cpp:650-661):
// continuous E^2 and discrete E are both exp(-w ds); phi = vf exp(-w ds) / Q_s.
template <typename T>
__global__ void solvent_phi_kernel(const T* w, double ds, double vf, const double* Qs, int M, double* out) {
    const double s = vf / *Qs;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x)
        out[i] = exp(-(double)w[i] * ds) * s;
}

// out = (accumulate ? out : 0) + sum_t scale_t src_t; blockIdx.y = r sums the sources
// src_t + r step_t into out + r out_step (replicated plans)
struct SumBatch {
    int n;
    int accumulate;
    const double* src[MAX_BATCH];
    double scale[MAX_BATCH];
    size_t step[MAX_BATCH];
    size_t out_step;
};

static __global__ void sum_kernel(SumBatch b, int M, double* out) {
    const size_t r = blockIdx.y;
    out += r * b.out_step;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        double acc = b.accumulate ? out[i] : 0.0;
        for (int t = 0; t < b.n; t++) acc += b.scale[t] * b.src[t][r * b.step[t] + i];
        out[i] = acc;
    }
}

// Initial slices (CpuComputationContinuous.cpp:206-298, CpuComputationDiscrete.cpp:260-445)
enum InitOp { INIT_ONE = 0, INIT_QINIT = 1, INIT_SUM = 2, INIT_PROD = 3 };
constexpr int INIT_MAX_SRC = 8;
constexpr int INIT_BATCH = 8;

template <typename T>
struct InitItem {
    T* dst;
    int op;
    int accumulate;   // SUM/PROD continue onto dst (deps beyond INIT_MAX_SRC)
    int n_src;
    const T* src[INIT_MAX_SRC];
    T coef[INIT_MAX_SRC];
    const double* qinit;
    const T* w;       // factor exp(-w*wcoef) after the op (nullptr: none)
    T wcoef;
    const T* mask;
};

template <typename T>
struct InitBatch {
    int n;
    InitItem<T> it[INIT_BATCH];
};

template <typename T>
__global__ void init_kernel(InitBatch<T> b, int M) {
    const InitItem<T>& it = b.it[blockIdx.y];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        T v;
        if (it.op == INIT_ONE) {
            v = (T)1;
        } else if (it.op == INIT_QINIT) {
            v = (T)it.qinit[i];
        } else if (it.op == INIT_SUM) {
            v = it.accumulate ? it.dst[i] : (T)0;
            for (int s = 0; s < it.n_src; s++) v += it.src[s][i] * it.coef[s];
        } else {
            v = it.accumulate ? it.dst[i] : (T)1;
            for (int s = 0; s < it.n_src; s++) v *= it.src[s][i];
        }
        if (it.w) v *= it.wcoef < (T)0 ? it.w[i] : fexp<T>(-it.w[i] * it.wcoef);
        if (it.mask) v *= it.mask[i];
        it.dst[i] = v;
    }
}

template <typename T>
__global__ void convert_kernel(const double* in, T* out, int M) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) out[i] = (T)in[i];
}

// out = exp(-w coef), the factor of every discrete step (w fixed during a call)
template <typename T>
__global__ void exp_table_kernel(const T* w, T coef, T* out, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = fexp<T>(-w[i] * coef);
}

// FP64 fields (one per blockIdx.y) into consecutive fields of the solver's precision
struct ConvertBatch {
    int n;
    const double* src[MAX_BATCH];
};

template <typename T>
__global__ void convert_batch_kernel(ConvertBatch b, T* out, int M) {
    const double* in = b.src[blockIdx.y];
    T* o = out + (size_t)blockIdx.y * M;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) o[i] = (T)in[i];
}

template <typename T>
__global__ void widen_kernel(const T* in, double* out, int M) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) out[i] = (double)in[i];
}

// Segment stress (CpuSolverPseudo.cpp:250-356): for term t with spectra A_t, B_t,
// sigma_d = sum_k bsq [B(k)] Re(A_k conj B_k) basis_d(k), basis_d = (2 pi k_d / L_d)^2
// times 2 for 0 < kz < nz/2 (Pseudo.cpp:84-153).
template <typename T>
struct StressItem {
    const typename vec2<T>::type* a;
    const typename vec2<T>::type* b;
    int table;      // Boltzmann table row or -1
    double bsq;
};

template <typename T>
struct StressBatch {
    int n;
    StressItem<T> it[MAX_BATCH];
};

struct StressGeom {
    int n0, n1, nz, nzc;
    int sp;              // row pitch of the spectra (nzc, or padded)
    double f0, f1, f2;   // (2 pi / L_d)^2 on the 3D-mapped axes
};

template <typename T>
__global__ void __launch_bounds__(RED_THREADS) stress_stage1(StressBatch<T> b, StressGeom g, const T* bx, const T* by,
                                                             const T* bz, double* partials) {
    __shared__ double ws[16];
    const StressItem<T>& it = b.it[blockIdx.y];
    const int Mc = g.n0 * g.n1 * g.nzc;
    double s0 = 0, s1 = 0, s2 = 0;
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < Mc; f += gridDim.x * blockDim.x) {
        const int k = f % g.nzc;
        const int r = f / g.nzc;
        const int j = r % g.n1;
        const int i = r / g.n1;
        const int it0 = i > g.n0 / 2 ? g.n0 - i : i;
        const int jt = j > g.n1 / 2 ? g.n1 - j : j;
        const size_t e = (size_t)r * g.sp + k;
        const auto A = it.a[e];
        const auto B = it.b[e];
        double coeff = it.bsq * ((double)A.x * (double)B.x + (double)A.y * (double)B.y);
        if (it.table >= 0)
            coeff *= (double)bx[(size_t)it.table * g.n0 + i] * (double)by[(size_t)it.table * g.n1 + j] *
                     (double)bz[(size_t)it.table * g.nzc + k];
        const double wgt = (k != 0 && 2 * k != g.nz) ? 2.0 : 1.0;
        s0 += coeff * (it0 * it0 * g.f0 * wgt);
        s1 += coeff * (jt * jt * g.f1 * wgt);
        s2 += coeff * (k * k * g.f2 * wgt);
    }
    s0 = block_sum(s0, ws);
    s1 = block_sum(s1, ws);
    s2 = block_sum(s2, ws);
    if (threadIdx.x == 0) {
        double* p = partials + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 3;
        p[0] = s0;
        p[1] = s1;
        p[2] = s2;
    }
}

static __global__ void __launch_bounds__(RED_BLOCKS) stress_stage2(const double* partials, double* out) {
    __shared__ double ws[16];
    for (int d = 0; d < 3; d++) {
        double v = partials[((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 3 + d];
        v = block_sum(v, ws);
        if (threadIdx.x == 0) out[blockIdx.x * 3 + d] = v;
    }
}

// The stress above with the x transforms of both spectra fused into it (wave-local N^3
// grids): workgroup (pencil j, kz tile; term p) takes the slab pass's output columns of A and
// of B, transforms each along x in turn and reduces its elements' terms of stress_stage1
// without writing either spectrum. Each spectrum is read once after the slab pass, not
// written by a pencil pass and read again. partials[(p gridDim.x + block) 3 + d].
template <typename T, int N, int NLW>
__global__ void __launch_bounds__(64) stress_pencil_wl(StressBatch<T> batch, StressGeom sg,
                                                       const typename vec2<T>::type* twx, const T* bx, const T* by,
                                                       const T* bz, double* partials) {
    using T2 = typename vec2<T>::type;
    using G = WPencil<N, NLW>;
    using F = typename G::F;
    constexpr int NL = G::NL, NZC = G::NZC;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T2* S = reinterpret_cast<T2*>(smem_raw);   // line l, element i at S[i NL + l]
    T2* Wx = S + N * NL;
    const int lane = threadIdx.x;
    const StressItem<T>& it = batch.it[blockIdx.y];
    const int j = blockIdx.x / G::TILES;
    const int k0 = (blockIdx.x - j * G::TILES) * NL;
    const int nl = min(NL, NZC - k0);
    constexpr int SP = wspec_pitch<T2>(N);
    constexpr size_t ROW = (size_t)N * SP;
    for (int m = lane; m < N; m += 64) Wx[m] = twx[m];
    using M0 = WStageMap<N, false, 0, NL, true>;
    using MF = WStageMap<N, false, F::n - 1, NL, true>;
    constexpr int R0 = F::r(0), RF = F::r(F::n - 1);
    auto load = [&](const T2* c, T2 (&v)[M0::NT][R0]) {
#pragma unroll
        for (int t = 0; t < M0::NT; t++) {
            const int g = lane + 64 * t;
            if (M0::ok(g, nl)) {
                const int l = M0::line(g), i0 = M0::idx(g);
#pragma unroll
                for (int r = 0; r < R0; r++) v[t][r] = c[(i0 + r * M0::B) * ROW + l];
            }
        }
    };
    // A's transformed columns stay in registers, at the positions where B's come out
    T2 fa[MF::NT][RF];
    {
        T2 v[M0::NT][R0];
        load(it.a + (size_t)j * SP + k0, v);
        wave_sync();
        wl_stages<-1, T, T2, N, false, NL, true, 1, NL, false, 0>(v, S, Wx, lane, nl, [&](auto& o) {
#pragma unroll
            for (int t = 0; t < MF::NT; t++)
#pragma unroll
                for (int r = 0; r < RF; r++) fa[t][r] = o[t][r];
        });
    }
    double s0 = 0, s1 = 0, s2 = 0;
    {
        T2 v[M0::NT][R0];
        load(it.b + (size_t)j * SP + k0, v);
        wave_sync();
        wl_stages<-1, T, T2, N, false, NL, true, 1, NL, false, 0>(v, S, Wx, lane, nl, [&](auto& o) {
            const int jt = j > N / 2 ? N - j : j;
#pragma unroll
            for (int t = 0; t < MF::NT; t++) {
                const int g = lane + 64 * t;
                if (MF::ok(g, nl)) {
                    const int k = k0 + MF::line(g), i0 = MF::idx(g);
                    const double wgt = (k != 0 && 2 * k != N) ? 2.0 : 1.0;
#pragma unroll
                    for (int r = 0; r < RF; r++) {
                        const int i = i0 + r * MF::B;
                        const int it0 = i > N / 2 ? N - i : i;
                        const T2 A = fa[t][r], B = o[t][r];
                        double coeff = it.bsq * ((double)A.x * (double)B.x + (double)A.y * (double)B.y);
                        if (it.table >= 0)
                            coeff *= (double)bx[(size_t)it.table * N + i] * (double)by[(size_t)it.table * N + j] *
                                     (double)bz[(size_t)it.table * NZC + k];
                        s0 += coeff * (it0 * it0 * sg.f0 * wgt);
                        s1 += coeff * (jt * jt * sg.f1 * wgt);
                        s2 += coeff * (k * k * sg.f2 * wgt);
                    }
                }
            }
        });
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if (lane == 0) {
        double* p = partials + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 3;
        p[0] = s0;
        p[1] = s1;
        p[2] = s2;
    }
}

// second stage of stress_pencil_wl: out[3 p + d] = the sum of term p's n partials, in a fixed order
static __global__ void __launch_bounds__(256) stress_stage2_n(const double* partials, int n, double* out) {
    __shared__ double ws[16];
    for (int d = 0; d < 3; d++) {
        double v = 0.0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) v += partials[((size_t)blockIdx.x * n + i) * 3 + d];
        v = block_sum(v, ws);
        if (threadIdx.x == 0) out[blockIdx.x * 3 + d] = v;
    }
}

static_assert(sizeof(PhiBatch<double>) <= 4000, "phi batch exceeds the kernel-argument budget");
static_assert(sizeof(StressBatch<double>) <= 4000, "stress batch exceeds the kernel-argument budget");
static_assert(sizeof(SumBatch) <= 4000, "sum batch exceeds the kernel-argument budget");

}  // namespace fts
