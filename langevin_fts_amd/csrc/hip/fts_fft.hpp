// LDS-staged mixed-radix Stockham FFT building blocks and the two 3D pass kernels
// of the pseudo-spectral propagator step, written for gfx950 (64-wide waves,
// 160 KiB LDS per CU).
//
// Layout in HBM (C order, as MklFFT3D.cpp:18-19): real field q[nx][ny][nz],
// half spectrum c[nx][ny][nzc], nzc = nz/2 + 1. 1D/2D grids use leading unit axes.
//
// A 3D round trip  q -> post * IFFT[ B * FFT[ pre * q ] ]  is three HBM passes:
//   slab pass    (one workgroup per x-slab): z real FFT + y FFT in LDS
//   pencil pass  (one workgroup per (y, kz-tile)): x FFT, * B(k), x IFFT in LDS
//   slab pass    : y IFFT + z real IFFT, * post, (then the next step's pre and
//                  forward z/y FFT in the same workgroup, so consecutive steps
//                  share one slab pass).
// The Boltzmann factor is separable, B(k) = Bx[i] By[j] Bz[k], so only three
// short 1D tables per (monomer, bond kind) are read instead of an Mc-sized table.
//
// Two real z-rows are transformed as one complex row (x_a + i x_b) and split
// with the Hermitian identities; this works for odd and even lengths alike.
#pragma once

#include "fts_common.hpp"

namespace fts {

constexpr int FFT_MAX_STAGES = 12;
constexpr int MAX_BATCH = 32;   // work items per launch (passed by value, kernel arguments < 4 KiB)
constexpr int FFT_MAXE = 8;     // complex values held per thread per FFT stage
constexpr int SPLIT_MAXE = 4;   // complex pairs held per thread in the real-row split/merge

struct AxisPlan {
    int L;
    int n_stages;
    int radix[FFT_MAX_STAGES];
    int group;   // lines transformed together (bounded by FFT_MAXE values per thread)
};

// radices outside {2, 3, 4, 5, 8} need the kernel variant with generic odd radices
inline bool plan_is_big(const AxisPlan& p) {
    for (int s = 0; s < p.n_stages; s++)
        if (p.radix[s] != 2 && p.radix[s] != 3 && p.radix[s] != 4 && p.radix[s] != 5 && p.radix[s] != 8) return true;
    return false;
}

// Radix schedule for one axis: 8s, then 4s, then at most one 2, then odd primes.
inline AxisPlan plan_axis(int L) {
    AxisPlan p{};
    p.L = L;
    int m = L;
    int n = 0;
    while (m % 8 == 0) { p.radix[n++] = 8; m /= 8; }
    while (m % 4 == 0) { p.radix[n++] = 4; m /= 4; }
    while (m % 2 == 0) { p.radix[n++] = 2; m /= 2; }
    for (int f = 3; m > 1; f += 2) {
        while (m % f == 0) {
            FTS_REQUIRE(f <= 31, "grid length " + std::to_string(L) + " has a prime factor above 31");
            FTS_REQUIRE(n < FFT_MAX_STAGES, "too many FFT stages");
            p.radix[n++] = f;
            m /= f;
        }
    }
    p.n_stages = n;
    return p;
}

// Butterflies a thread holds in a stage of radix R.
inline int stage_slots(int R) { return FFT_MAXE / R > 0 ? FFT_MAXE / R : 1; }

// Number of lines of plan p that one pass of the stages can hold at block size bd.
inline int group_lines(const AxisPlan& p, int bd) {
    int g = 1 << 30;
    for (int s = 0; s < p.n_stages; s++) {
        const int R = p.radix[s];
        g = std::min(g, stage_slots(R) * bd / (p.L / R));
    }
    FTS_REQUIRE(g >= 1, "FFT line of length " + std::to_string(p.L) + " does not fit the block");
    return g;
}

// ------------------------------------------------------------------ complex helpers
template <typename T2>
__device__ __forceinline__ T2 cadd(T2 a, T2 b) { return T2{a.x + b.x, a.y + b.y}; }
template <typename T2>
__device__ __forceinline__ T2 csub(T2 a, T2 b) { return T2{a.x - b.x, a.y - b.y}; }
template <typename T2>
__device__ __forceinline__ T2 cmul(T2 a, T2 b) { return T2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
// multiply by i*s (s = +1 or -1)
template <int S, typename T2>
__device__ __forceinline__ T2 mul_is(T2 a) {
    if constexpr (S > 0) return T2{-a.y, a.x};
    else return T2{a.y, -a.x};
}
template <typename T2, typename T>
__device__ __forceinline__ T2 cscale(T2 a, T s) { return T2{a.x * s, a.y * s}; }

// root of unity W_L^m with direction DIR (-1 forward, +1 inverse); table holds forward roots
template <int DIR, typename T2>
__device__ __forceinline__ T2 root(const T2* W, int m) {
    T2 w = W[m];
    if constexpr (DIR > 0) w.y = -w.y;
    return w;
}

// ------------------------------------------------------------------ butterflies
template <int R, int DIR, typename T, typename T2>
struct Dft;

template <int DIR, typename T, typename T2>
struct Dft<2, DIR, T, T2> {
    __device__ static void run(T2* v, const T2*, int) {
        T2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int DIR, typename T, typename T2>
struct Dft<4, DIR, T, T2> {
    __device__ static void run(T2* v, const T2*, int) {
        T2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        T2 t2 = cadd(v[1], v[3]), t3 = mul_is<DIR>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[1] = cadd(t1, t3);
        v[2] = csub(t0, t2);
        v[3] = csub(t1, t3);
    }
};

template <int DIR, typename T, typename T2>
struct Dft<8, DIR, T, T2> {
    __device__ static void run(T2* v, const T2*, int) {
        const T h = (T)0.70710678118654752440084436210485;
        T2 e[4] = {v[0], v[2], v[4], v[6]};
        T2 o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, DIR, T, T2>::run(e, nullptr, 0);
        Dft<4, DIR, T, T2>::run(o, nullptr, 0);
        // o[k] *= W8^k, W8 = (1 + DIR i) h
        T2 w1 = T2{h, (T)DIR * h};
        o[1] = cmul(o[1], w1);
        o[2] = mul_is<DIR>(o[2]);
        T2 w3 = T2{-h, (T)DIR * h};
        o[3] = cmul(o[3], w3);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = cadd(e[k], o[k]);
            v[k + 4] = csub(e[k], o[k]);
        }
    }
};

template <int DIR, typename T, typename T2>
struct Dft<3, DIR, T, T2> {
    __device__ static void run(T2* v, const T2*, int) {
        const T s = (T)0.86602540378443864676372317075294;
        T2 u = cadd(v[1], v[2]);
        T2 d = csub(v[1], v[2]);
        T2 t = T2{v[0].x - (T)0.5 * u.x, v[0].y - (T)0.5 * u.y};
        T2 r = cscale(mul_is<DIR>(d), s);
        v[0] = cadd(v[0], u);
        v[1] = cadd(t, r);
        v[2] = csub(t, r);
    }
};

template <int DIR, typename T, typename T2>
struct Dft<5, DIR, T, T2> {
    __device__ static void run(T2* v, const T2*, int) {
        const T c1 = (T)0.30901699437494742410229341718282;   // cos(2 pi/5)
        const T c2 = (T)-0.80901699437494742410229341718282;  // cos(4 pi/5)
        const T s1 = (T)0.95105651629515357211643933337938;   // sin(2 pi/5)
        const T s2 = (T)0.58778525229247312916870595463907;   // sin(4 pi/5)
        T2 b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
        T2 d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
        T2 a0 = v[0];
        T2 t1 = T2{a0.x + c1 * b1.x + c2 * b2.x, a0.y + c1 * b1.y + c2 * b2.y};
        T2 t2 = T2{a0.x + c2 * b1.x + c1 * b2.x, a0.y + c2 * b1.y + c1 * b2.y};
        T2 u1 = mul_is<DIR>(T2{s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y});
        T2 u2 = mul_is<DIR>(T2{s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y});
        v[0] = cadd(a0, cadd(b1, b2));
        v[1] = cadd(t1, u1);
        v[4] = csub(t1, u1);
        v[2] = cadd(t2, u2);
        v[3] = csub(t2, u2);
    }
};

// Generic odd radix: y_q = sum_r v_r W_R^{rq}, W_R^m = W_L^{m L/R} from the LDS table.
template <int R, int DIR, typename T, typename T2>
struct Dft {
    __device__ static void run(T2* v, const T2* W, int L) {
        T2 y[R];
        const int st = L / R;
#pragma unroll
        for (int q = 0; q < R; q++) {
            T2 s = v[0];
#pragma unroll
            for (int r = 1; r < R; r++) s = cadd(s, cmul(v[r], root<DIR>(W, ((r * q) % R) * st)));
            y[q] = s;
        }
#pragma unroll
        for (int q = 0; q < R; q++) v[q] = y[q];
    }
};

// One Stockham stage over n_lines transforms of length L stored in LDS at
// S[line*line_pitch + idx*elem_stride]. Ns = product of the radices done so far.
// All loads of the stage complete before any store (barrier in between), so the
// stage runs in place.
template <int R, int DIR, typename T, typename T2>
__device__ __forceinline__ void fft_stage(T2* S, int n_lines, int line_pitch, int elem_stride, int L, int Ns,
                                          const T2* W) {
    constexpr int NB = FFT_MAXE / R > 0 ? FFT_MAXE / R : 1;
    const int LR = L / R;
    const int total = n_lines * LR;
    const int tw_step = L / (Ns * R);
    T2 v[NB][R];
#pragma unroll
    for (int t = 0; t < NB; t++) {
        const int g = threadIdx.x + t * blockDim.x;
        if (g < total) {
            const int line = g / LR;
            const int j = g - line * LR;
            const T2* base = S + line * line_pitch;
            const int k = j % Ns;
#pragma unroll
            for (int r = 0; r < R; r++) v[t][r] = base[(j + r * LR) * elem_stride];
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) v[t][r] = cmul(v[t][r], root<DIR>(W, (r * k * tw_step) % L));
            }
            Dft<R, DIR, T, T2>::run(v[t], W, L);
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NB; t++) {
        const int g = threadIdx.x + t * blockDim.x;
        if (g < total) {
            const int line = g / LR;
            const int j = g - line * LR;
            T2* base = S + line * line_pitch;
            const int k = j % Ns;
            const int d = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) base[(d + r * Ns) * elem_stride] = v[t][r];
        }
    }
    __syncthreads();
}

// Full transform of n_lines lines (caller has synchronised LDS before the call).
// Lines are independent, so they are processed in groups of p.group lines, which
// bounds the registers a thread holds whatever the slab size. BIG selects the
// kernel variant that also carries the generic odd radices 7..31.
template <int DIR, bool BIG, typename T, typename T2>
__device__ void lds_fft(T2* S, int n_lines, int line_pitch, int elem_stride, const AxisPlan& p, const T2* W) {
    if (p.n_stages == 0) return;
    for (int l0 = 0; l0 < n_lines; l0 += p.group) {
        const int cnt = min(p.group, n_lines - l0);
        T2* Sg = S + l0 * line_pitch;
        int Ns = 1;
        for (int s = 0; s < p.n_stages; s++) {
            const int R = p.radix[s];
            switch (R) {
                case 2: fft_stage<2, DIR, T>(Sg, cnt, line_pitch, elem_stride, p.L, Ns, W); break;
                case 3: fft_stage<3, DIR, T>(Sg, cnt, line_pitch, elem_stride, p.L, Ns, W); break;
                case 4: fft_stage<4, DIR, T>(Sg, cnt, line_pitch, elem_stride, p.L, Ns, W); break;
                case 5: fft_stage<5, DIR, T>(Sg, cnt, line_pitch, elem_stride, p.L, Ns, W); break;
                case 8: fft_stage<8, DIR, T>(Sg, cnt, line_pitch, elem_stride, p.L, Ns, W); break;
                default:
                    if constexpr (BIG) {
                        switch (R) {
                            case 7: fft_stage<7, DIR, T>(Sg, cnt, line_pitch, elem_stride, p.L, Ns, W); break;
                            case 11: fft_stage<11, DIR, T>(Sg, cnt, line_pitch, elem_stride, p.L, Ns, W); break;
                            case 13: fft_stage<13, DIR, T>(Sg, cnt, line_pitch, elem_stride, p.L, Ns, W); break;
                            case 17: fft_stage<17, DIR, T>(Sg, cnt, line_pitch, elem_stride, p.L, Ns, W); break;
                            case 19: fft_stage<19, DIR, T>(Sg, cnt, line_pitch, elem_stride, p.L, Ns, W); break;
                            case 23: fft_stage<23, DIR, T>(Sg, cnt, line_pitch, elem_stride, p.L, Ns, W); break;
                            case 29: fft_stage<29, DIR, T>(Sg, cnt, line_pitch, elem_stride, p.L, Ns, W); break;
                            case 31: fft_stage<31, DIR, T>(Sg, cnt, line_pitch, elem_stride, p.L, Ns, W); break;
                            default: break;
                        }
                    }
                    break;
            }
            Ns *= R;
        }
    }
}


// ------------------------------------------------------------------ slab pass
struct SlabGeom {
    int ny, nz, nzc;
    int P;       // LDS row pitch in complex elements (>= nzc)
    int rows;    // 2 * ceil(ny / 2)
    AxisPlan ay, az;
    double inv_m;
};

template <typename T>
struct SlabItem {
    const T* qin;      // real input of the forward part (FWD mode)
    T* qout;           // real output of the inverse part (nullptr: no write)
    typename vec2<T>::type* c0;
    typename vec2<T>::type* c1;
    const T* w;        // field for exp(-w*coef) factors (nullptr: all factors 1; coef < 0: w is the factor)
    const T* mask;     // nullptr: no mask
    T post0, post1;    // inverse part: r_o *= exp(-w*post_o) / M
    T pre0, pre1;      // forward part: spectrum_o = FFT[exp(-w*pre_o) q]
    int n_out;         // FWD mode: 1 or 2 spectra
    int forward;       // INV modes: continue with the forward part
};

template <typename T>
struct SlabBatch {
    int n;
    SlabItem<T> it[MAX_BATCH];
};

enum SlabMode { SLAB_FWD = 0, SLAB_INV1 = 1, SLAB_INV2 = 2 };

template <typename T>
__device__ __forceinline__ T fexp(T x);
template <>
__device__ __forceinline__ double fexp<double>(double x) { return exp(x); }
template <>
__device__ __forceinline__ float fexp<float>(float x) { return __expf(x); }

// In LDS the pair of rows (2p, 2p+1) of a slab is packed as one complex row of
// length nz starting at S[2p*P]: row 2p in .x, row 2p+1 in .y. Each thread owns
// the real slab elements e = tid + t*blockDim (t < MAXR), e = j*nz + k.
template <typename T2>
__device__ __forceinline__ T2* packed_slot(T2* S, int P, int j, int k) {
    return S + (j & ~1) * P + k;
}

// packed real rows -> half spectra S[j][k] (z forward + split), then y forward
template <bool BIG, typename T, typename T2>
__device__ void slab_forward_lds(T2* S, const SlabGeom& g, const T2* Wy, const T2* Wz) {
    const int np = g.rows >> 1;
    lds_fft<-1, BIG, T>(S, np, 2 * g.P, 1, g.az, Wz);
    // split: X_a[k] = (Z[k] + conj Z[-k]) / 2, X_b[k] = -i (Z[k] - conj Z[-k]) / 2;
    // pairs are independent, processed in groups that fit FFT_MAXE values per thread
    const int pg = max(1, (SPLIT_MAXE * (int)blockDim.x) / g.nzc);
    for (int p0 = 0; p0 < np; p0 += pg) {
        const int total = min(pg, np - p0) * g.nzc;
        T2 va[SPLIT_MAXE], vb[SPLIT_MAXE];
#pragma unroll
        for (int t = 0; t < SPLIT_MAXE; t++) {
            const int f = threadIdx.x + t * blockDim.x;
            if (f < total) {
                const int p = p0 + f / g.nzc, k = f % g.nzc;
                const T2* Z = S + 2 * p * g.P;
                const T2 A = Z[k];
                const int km = k == 0 ? 0 : g.nz - k;
                const T2 B = T2{Z[km].x, -Z[km].y};
                va[t] = T2{(T)0.5 * (A.x + B.x), (T)0.5 * (A.y + B.y)};
                vb[t] = T2{(T)0.5 * (A.y - B.y), (T)-0.5 * (A.x - B.x)};
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < SPLIT_MAXE; t++) {
            const int f = threadIdx.x + t * blockDim.x;
            if (f < total) {
                const int p = p0 + f / g.nzc, k = f % g.nzc;
                S[2 * p * g.P + k] = va[t];
                S[(2 * p + 1) * g.P + k] = vb[t];
            }
        }
        __syncthreads();
    }
    lds_fft<-1, BIG, T>(S, g.nzc, 1, g.P, g.ay, Wy);
}

// half spectra S[j][k] -> y inverse, merge row pairs, z inverse (unnormalised)
template <bool BIG, typename T, typename T2>
__device__ void slab_inverse_lds(T2* S, const SlabGeom& g, const T2* Wy, const T2* Wz) {
    lds_fft<+1, BIG, T>(S, g.nzc, 1, g.P, g.ay, Wy);
    const int np = g.rows >> 1;
    const int pg = max(1, (SPLIT_MAXE * (int)blockDim.x) / g.nz);
    for (int p0 = 0; p0 < np; p0 += pg) {
        const int total = min(pg, np - p0) * g.nz;
        T2 vz[SPLIT_MAXE];
#pragma unroll
        for (int t = 0; t < SPLIT_MAXE; t++) {
            const int f = threadIdx.x + t * blockDim.x;
            if (f < total) {
                const int p = p0 + f / g.nz, k = f % g.nz;
                const bool upper = k >= g.nzc;
                const int kk = upper ? g.nz - k : k;
                T2 a = S[2 * p * g.P + kk];
                T2 b = S[(2 * p + 1) * g.P + kk];
                if (kk == 0 || 2 * kk == g.nz) {
                    a.y = 0;
                    b.y = 0;
                }
                if (upper) {
                    a.y = -a.y;
                    b.y = -b.y;
                }
                // Z = X_a + i X_b
                vz[t] = T2{a.x - b.y, a.y + b.x};
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < SPLIT_MAXE; t++) {
            const int f = threadIdx.x + t * blockDim.x;
            if (f < total) {
                const int p = p0 + f / g.nz, k = f % g.nz;
                S[2 * p * g.P + k] = vz[t];
            }
        }
        __syncthreads();
    }
    lds_fft<+1, BIG, T>(S, np, 2 * g.P, 1, g.az, Wz);
}

template <typename T2>
__device__ __forceinline__ void load_spectrum(T2* S, const SlabGeom& g, const T2* c) {
    const int total = g.ny * g.nzc;
    for (int f = threadIdx.x; f < total; f += blockDim.x) {
        const int j = f / g.nzc, k = f - j * g.nzc;
        S[j * g.P + k] = c[f];
    }
    // odd ny: the partner row of the last row is merged into the same complex
    // z-row by the inverse; it must hold zeros, not what the LDS held before
    if (g.ny & 1)
        for (int k = threadIdx.x; k < g.nzc; k += blockDim.x) S[g.ny * g.P + k] = T2{0, 0};
}

template <typename T2>
__device__ __forceinline__ void store_spectrum(const T2* S, const SlabGeom& g, T2* c) {
    const int total = g.ny * g.nzc;
    for (int f = threadIdx.x; f < total; f += blockDim.x) {
        const int j = f / g.nzc, k = f - j * g.nzc;
        c[f] = S[j * g.P + k];
    }
}

template <typename T>
__device__ __forceinline__ T factor(T w, bool has_w, T coef) {
    // coef < 0: the array already holds the factor exp(-w ds) (a table made once per call)
    return (has_w && coef != (T)0) ? (coef < (T)0 ? w : fexp<T>(-w * coef)) : (T)1;
}

// write this thread's real values times exp(-w*coef) into the packed rows
template <int MAXR, typename T, typename T2>
__device__ __forceinline__ void pack_real(T2* S, const SlabGeom& g, const T* q, const T* wv, bool has_w, T coef) {
    const int nreal = g.ny * g.nz;
#pragma unroll
    for (int t = 0; t < MAXR; t++) {
        const int e = threadIdx.x + t * blockDim.x;
        if (e < nreal) {
            const int j = e / g.nz, k = e - j * g.nz;
            T2* s = packed_slot(S, g.P, j, k);
            const T v = q[t] * factor<T>(wv[t], has_w, coef);
            if (j & 1) s->y = v;
            else s->x = v;
        }
    }
    if (g.ny & 1) {  // phantom partner of the last row
        for (int k = threadIdx.x; k < g.nz; k += blockDim.x) S[(g.ny - 1) * g.P + k].y = 0;
    }
    __syncthreads();
}

template <typename T, int MODE, int MAXR, bool BIG>
__global__ void __launch_bounds__(MAXR == 8 ? 512 : 1024) slab_kernel(SlabGeom g, SlabBatch<T> batch, const typename vec2<T>::type* twy,
                                                   const typename vec2<T>::type* twz) {
    using T2 = typename vec2<T>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T2* S = reinterpret_cast<T2*>(smem_raw);
    T2* Wy = S + g.rows * g.P;
    T2* Wz = Wy + g.ny;
    for (int m = threadIdx.x; m < g.ny; m += blockDim.x) Wy[m] = twy[m];
    for (int m = threadIdx.x; m < g.nz; m += blockDim.x) Wz[m] = twz[m];

    const SlabItem<T>& it = batch.it[blockIdx.y];
    const int i = blockIdx.x;
    const int nreal = g.ny * g.nz;
    const size_t roff = (size_t)i * nreal;
    const size_t coff = (size_t)i * g.ny * g.nzc;
    const T inv_m = (T)g.inv_m;
    const bool has_w = it.w != nullptr;

    T q[MAXR];
    T wv[MAXR];
#pragma unroll
    for (int t = 0; t < MAXR; t++) {
        const int e = threadIdx.x + t * blockDim.x;
        wv[t] = (has_w && e < nreal) ? it.w[roff + e] : (T)0;
        q[t] = (T)0;
    }
    bool do_fwd = true;
    if constexpr (MODE == SLAB_FWD) {
#pragma unroll
        for (int t = 0; t < MAXR; t++) {
            const int e = threadIdx.x + t * blockDim.x;
            if (e < nreal) q[t] = it.qin[roff + e];
        }
    } else {
        load_spectrum(S, g, it.c0 + coff);
        __syncthreads();
        slab_inverse_lds<BIG, T>(S, g, Wy, Wz);
        // q = r0 exp(-w post0) / M
#pragma unroll
        for (int t = 0; t < MAXR; t++) {
            const int e = threadIdx.x + t * blockDim.x;
            if (e < nreal) {
                const int j = e / g.nz, k = e - j * g.nz;
                const T2 s = *packed_slot(S, g.P, j, k);
                q[t] = ((j & 1) ? s.y : s.x) * factor<T>(wv[t], has_w, it.post0) * inv_m;
            }
        }
        if constexpr (MODE == SLAB_INV2) {
            __syncthreads();
            load_spectrum(S, g, it.c1 + coff);
            __syncthreads();
            slab_inverse_lds<BIG, T>(S, g, Wy, Wz);
            // 4th-order Richardson combination (CpuSolverPseudo.cpp:179-180)
#pragma unroll
            for (int t = 0; t < MAXR; t++) {
                const int e = threadIdx.x + t * blockDim.x;
                if (e < nreal) {
                    const int j = e / g.nz, k = e - j * g.nz;
                    const T2 s = *packed_slot(S, g.P, j, k);
                    const T q2 = ((j & 1) ? s.y : s.x) * factor<T>(wv[t], has_w, it.post1) * inv_m;
                    q[t] = ((T)4.0 * q2 - q[t]) / (T)3.0;
                }
            }
        }
        if (it.mask) {
#pragma unroll
            for (int t = 0; t < MAXR; t++) {
                const int e = threadIdx.x + t * blockDim.x;
                if (e < nreal) q[t] *= it.mask[roff + e];
            }
        }
        if (it.qout) {
#pragma unroll
            for (int t = 0; t < MAXR; t++) {
                const int e = threadIdx.x + t * blockDim.x;
                if (e < nreal) it.qout[roff + e] = q[t];
            }
        }
        do_fwd = it.forward != 0;
        __syncthreads();
    }
    if (!do_fwd) return;

    const int n_out = (MODE == SLAB_FWD) ? it.n_out : (MODE == SLAB_INV2 ? 2 : 1);
    for (int o = 0; o < n_out; o++) {
        if (o > 0) __syncthreads();
        pack_real<MAXR, T>(S, g, q, wv, has_w, o == 0 ? it.pre0 : it.pre1);
        slab_forward_lds<BIG, T>(S, g, Wy, Wz);
        store_spectrum(S, g, (o == 0 ? it.c0 : it.c1) + coff);
    }
}

// ------------------------------------------------------------------ pencil (x) pass
struct XGeom {
    int nx, ny, nzc;
    int C;         // kz tile width
    int n_tiles;   // ceil(nzc / C)
    int P;         // LDS pitch (>= C)
    AxisPlan ax;
};

template <typename T>
struct XItem {
    typename vec2<T>::type* c;
    int table;   // Boltzmann table row; -1: forward only; -2: inverse only
};

template <typename T>
struct XBatch {
    int n;
    XItem<T> it[MAX_BATCH];
};

// bx: [n_tables][nx], by: [n_tables][ny], bz: [n_tables][nzc]
template <typename T, bool BIG>
__global__ void __launch_bounds__(1024) xpass_kernel(XGeom g, XBatch<T> batch, const typename vec2<T>::type* twx,
                                                    const T* bx, const T* by, const T* bz) {
    using T2 = typename vec2<T>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T2* S = reinterpret_cast<T2*>(smem_raw);
    T2* Wx = S + g.nx * g.P;
    for (int m = threadIdx.x; m < g.nx; m += blockDim.x) Wx[m] = twx[m];

    const XItem<T>& it = batch.it[blockIdx.y];
    const int j = blockIdx.x / g.n_tiles;
    const int tile = blockIdx.x - j * g.n_tiles;
    const int k0 = tile * g.C;
    const int cw = min(g.C, g.nzc - k0);
    const size_t row = (size_t)g.ny * g.nzc;
    T2* c = it.c + (size_t)j * g.nzc + k0;

    const int total = g.nx * cw;
    for (int f = threadIdx.x; f < total; f += blockDim.x) {
        const int i = f / cw, k = f - i * cw;
        S[i * g.P + k] = c[i * row + k];
    }
    __syncthreads();
    if (it.table != -2) lds_fft<-1, BIG, T>(S, cw, 1, g.P, g.ax, Wx);
    if (it.table >= 0) {
        const T* tx = bx + (size_t)it.table * g.nx;
        const T byj = by[(size_t)it.table * g.ny + j];
        const T* tz = bz + (size_t)it.table * g.nzc + k0;
        for (int f = threadIdx.x; f < total; f += blockDim.x) {
            const int i = f / cw, k = f - i * cw;
            const T b = tx[i] * byj * tz[k];
            T2 v = S[i * g.P + k];
            S[i * g.P + k] = T2{v.x * b, v.y * b};
        }
        __syncthreads();
    }
    if (it.table != -1) lds_fft<+1, BIG, T>(S, cw, 1, g.P, g.ax, Wx);
    for (int f = threadIdx.x; f < total; f += blockDim.x) {
        const int i = f / cw, k = f - i * cw;
        c[i * row + k] = S[i * g.P + k];
    }
}

}  // namespace fts
