// Compile-time specialised slab and pencil kernels for cubic grids
// N = 32, 48, 64, 96, 128: the radix plan, line strides, thread mapping and
// per-thread element counts are template constants, so every index is a
// constant-folded multiply/shift and every loop is fully unrolled. Same data
// layout, same passes and same arithmetic as the runtime kernels of fts_fft.hpp.
#pragma once

#include "fts_fft.hpp"

namespace fts {

struct CPlan {
    int n;
    int r[8];
};

constexpr CPlan make_cplan(int L) {
    CPlan p{0, {0, 0, 0, 0, 0, 0, 0, 0}};
    int m = L;
    while (m % 8 == 0) { p.r[p.n++] = 8; m /= 8; }
    while (m % 4 == 0) { p.r[p.n++] = 4; m /= 4; }
    while (m % 2 == 0) { p.r[p.n++] = 2; m /= 2; }
    while (m % 3 == 0) { p.r[p.n++] = 3; m /= 3; }
    while (m % 5 == 0) { p.r[p.n++] = 5; m /= 5; }
    return p;
}

// One Stockham stage with every size a constant (same indexing as fft_stage).
template <int R, int DIR, typename T, typename T2, int L, int NS, int NLINES, int LP, int ES, int BD>
__device__ __forceinline__ void ct_stage(T2* S, const T2* W) {
    constexpr int LR = L / R;
    constexpr int TOTAL = NLINES * LR;
    constexpr int NB = (TOTAL + BD - 1) / BD;
    constexpr int TW = L / (NS * R);
    T2 v[NB][R];
#pragma unroll
    for (int t = 0; t < NB; t++) {
        const int g = threadIdx.x + t * BD;
        if (TOTAL % BD == 0 || g < TOTAL) {
            const int line = g / LR;
            const int j = g - line * LR;
            const T2* base = S + line * LP;
            const int k = j % NS;
#pragma unroll
            for (int r = 0; r < R; r++) v[t][r] = base[(j + r * LR) * ES];
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) v[t][r] = cmul(v[t][r], root<DIR>(W, (r * k * TW) % L));
            }
            Dft<R, DIR, T, T2>::run(v[t], W, L);
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NB; t++) {
        const int g = threadIdx.x + t * BD;
        if (TOTAL % BD == 0 || g < TOTAL) {
            const int line = g / LR;
            const int j = g - line * LR;
            T2* base = S + line * LP;
            const int k = j % NS;
            const int d = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) base[(d + r * NS) * ES] = v[t][r];
        }
    }
    __syncthreads();
}

template <int DIR, typename T, typename T2, int L, int NLINES, int LP, int ES, int BD, int S = 0, int NS = 1>
__device__ __forceinline__ void ct_fft_lines(T2* X, const T2* W) {
    constexpr CPlan P = make_cplan(L);
    if constexpr (S < P.n) {
        constexpr int R = P.r[S];
        ct_stage<R, DIR, T, T2, L, NS, NLINES, LP, ES, BD>(X, W);
        ct_fft_lines<DIR, T, T2, L, NLINES, LP, ES, BD, S + 1, NS * R>(X, W);
    }
}

// lines in groups of about 8 values per thread (bounded registers for any N)
template <int DIR, typename T, typename T2, int L, int NLINES, int LP, int ES, int BD, int G0 = 0>
__device__ __forceinline__ void ct_fft(T2* X, const T2* W) {
    constexpr int GL0 = (8 * BD) / L;
    constexpr int GL = GL0 < 1 ? 1 : (GL0 > NLINES ? NLINES : GL0);
    if constexpr (G0 < NLINES) {
        constexpr int CNT = (NLINES - G0) < GL ? (NLINES - G0) : GL;
        ct_fft_lines<DIR, T, T2, L, CNT, LP, ES, BD>(X + G0 * LP, W);
        ct_fft<DIR, T, T2, L, NLINES, LP, ES, BD, G0 + GL>(X, W);
    }
}

// ------------------------------------------------------------------ slab kernel, N^3
template <int N>
struct SlabShape {
    static constexpr int NZC = N / 2 + 1;
    static constexpr int P = NZC;            // LDS row pitch (complex)
    static constexpr int NP = N / 2;         // row pairs (N even)
    static constexpr int NREAL = N * N;
    static constexpr int BD = N <= 32 ? 256 : (N <= 48 ? 384 : (N <= 64 ? 512 : 1024));
    static constexpr int MAXR = (NREAL + BD - 1) / BD;
    static constexpr int NSPEC = N * NZC;
    static constexpr size_t lds_elems = (size_t)N * P + 2 * N;
};

template <typename T, typename T2, int N>
__device__ __forceinline__ void ct_slab_forward(T2* S, const T2* Wy, const T2* Wz) {
    using G = SlabShape<N>;
    ct_fft<-1, T, T2, N, G::NP, 2 * G::P, 1, G::BD>(S, Wz);
    // split Z[k] (rows 2p + i 2p+1) into the two half spectra
    constexpr int TOTAL = G::NP * G::NZC;
    constexpr int NT = (TOTAL + G::BD - 1) / G::BD;
    T2 va[NT], vb[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int f = threadIdx.x + t * G::BD;
        if (TOTAL % G::BD == 0 || f < TOTAL) {
            const int p = f / G::NZC, k = f - p * G::NZC;
            const T2* Z = S + 2 * p * G::P;
            const T2 A = Z[k];
            const int km = k == 0 ? 0 : N - k;
            const T2 B = T2{Z[km].x, -Z[km].y};
            va[t] = T2{(T)0.5 * (A.x + B.x), (T)0.5 * (A.y + B.y)};
            vb[t] = T2{(T)0.5 * (A.y - B.y), (T)-0.5 * (A.x - B.x)};
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int f = threadIdx.x + t * G::BD;
        if (TOTAL % G::BD == 0 || f < TOTAL) {
            const int p = f / G::NZC, k = f - p * G::NZC;
            S[2 * p * G::P + k] = va[t];
            S[(2 * p + 1) * G::P + k] = vb[t];
        }
    }
    __syncthreads();
    ct_fft<-1, T, T2, N, G::NZC, 1, G::P, G::BD>(S, Wy);
}

template <typename T, typename T2, int N>
__device__ __forceinline__ void ct_slab_inverse(T2* S, const T2* Wy, const T2* Wz) {
    using G = SlabShape<N>;
    ct_fft<+1, T, T2, N, G::NZC, 1, G::P, G::BD>(S, Wy);
    constexpr int TOTAL = G::NP * N;
    constexpr int NT = (TOTAL + G::BD - 1) / G::BD;
    T2 vz[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int f = threadIdx.x + t * G::BD;
        if (TOTAL % G::BD == 0 || f < TOTAL) {
            const int p = f / N, k = f - p * N;
            const bool upper = k >= G::NZC;
            const int kk = upper ? N - k : k;
            T2 a = S[2 * p * G::P + kk];
            T2 b = S[(2 * p + 1) * G::P + kk];
            if (kk == 0 || 2 * kk == N) {
                a.y = 0;
                b.y = 0;
            }
            if (upper) {
                a.y = -a.y;
                b.y = -b.y;
            }
            vz[t] = T2{a.x - b.y, a.y + b.x};
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int f = threadIdx.x + t * G::BD;
        if (TOTAL % G::BD == 0 || f < TOTAL) {
            const int p = f / N, k = f - p * N;
            S[2 * p * G::P + k] = vz[t];
        }
    }
    __syncthreads();
    ct_fft<+1, T, T2, N, G::NP, 2 * G::P, 1, G::BD>(S, Wz);
}

template <typename T, int MODE, int N>
__global__ void __launch_bounds__(SlabShape<N>::BD) slab_kernel_ct(SlabBatch<T> batch, const typename vec2<T>::type* twy,
                                                                  const typename vec2<T>::type* twz, double inv_m_d) {
    using T2 = typename vec2<T>::type;
    using G = SlabShape<N>;
    constexpr int BD = G::BD, MAXR = G::MAXR, NREAL = G::NREAL;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T2* S = reinterpret_cast<T2*>(smem_raw);
    T2* Wy = S + N * G::P;
    T2* Wz = Wy + N;
    for (int m = threadIdx.x; m < N; m += BD) {
        Wy[m] = twy[m];
        Wz[m] = twz[m];
    }
    const SlabItem<T>& it = batch.it[blockIdx.y];
    const size_t roff = (size_t)blockIdx.x * NREAL;
    const size_t coff = (size_t)blockIdx.x * G::NSPEC;
    const T inv_m = (T)inv_m_d;
    const bool has_w = it.w != nullptr;
    auto inb = [](int e) { return NREAL % BD == 0 || e < NREAL; };

    T q[MAXR];
    T wv[MAXR];
#pragma unroll
    for (int t = 0; t < MAXR; t++) {
        const int e = threadIdx.x + t * BD;
        wv[t] = (has_w && inb(e)) ? it.w[roff + e] : (T)0;
        q[t] = (T)0;
    }
    bool do_fwd = true;
    if constexpr (MODE == SLAB_FWD) {
#pragma unroll
        for (int t = 0; t < MAXR; t++) {
            const int e = threadIdx.x + t * BD;
            if (inb(e)) q[t] = it.qin[roff + e];
        }
    } else {
        const T2* c0 = it.c0 + coff;
        for (int f = threadIdx.x; f < G::NSPEC; f += BD) S[(f / G::NZC) * G::P + f % G::NZC] = c0[f];
        __syncthreads();
        ct_slab_inverse<T, T2, N>(S, Wy, Wz);
#pragma unroll
        for (int t = 0; t < MAXR; t++) {
            const int e = threadIdx.x + t * BD;
            if (inb(e)) {
                const int j = e / N, k = e - j * N;
                const T2 s = S[(j & ~1) * G::P + k];
                q[t] = ((j & 1) ? s.y : s.x) * factor<T>(wv[t], has_w, it.post0) * inv_m;
            }
        }
        if constexpr (MODE == SLAB_INV2) {
            __syncthreads();
            const T2* c1 = it.c1 + coff;
            for (int f = threadIdx.x; f < G::NSPEC; f += BD) S[(f / G::NZC) * G::P + f % G::NZC] = c1[f];
            __syncthreads();
            ct_slab_inverse<T, T2, N>(S, Wy, Wz);
#pragma unroll
            for (int t = 0; t < MAXR; t++) {
                const int e = threadIdx.x + t * BD;
                if (inb(e)) {
                    const int j = e / N, k = e - j * N;
                    const T2 s = S[(j & ~1) * G::P + k];
                    const T q2 = ((j & 1) ? s.y : s.x) * factor<T>(wv[t], has_w, it.post1) * inv_m;
                    q[t] = ((T)4.0 * q2 - q[t]) / (T)3.0;
                }
            }
        }
        if (it.mask) {
#pragma unroll
            for (int t = 0; t < MAXR; t++) {
                const int e = threadIdx.x + t * BD;
                if (inb(e)) q[t] *= it.mask[roff + e];
            }
        }
        if (it.qout) {
#pragma unroll
            for (int t = 0; t < MAXR; t++) {
                const int e = threadIdx.x + t * BD;
                if (inb(e)) it.qout[roff + e] = q[t];
            }
        }
        do_fwd = it.forward != 0;
        __syncthreads();
    }
    if (!do_fwd) return;
    const int n_out = (MODE == SLAB_FWD) ? it.n_out : (MODE == SLAB_INV2 ? 2 : 1);
    for (int o = 0; o < n_out; o++) {
        if (o > 0) __syncthreads();
        const T coef = o == 0 ? it.pre0 : it.pre1;
#pragma unroll
        for (int t = 0; t < MAXR; t++) {
            const int e = threadIdx.x + t * BD;
            if (inb(e)) {
                const int j = e / N, k = e - j * N;
                T2* s = S + (j & ~1) * G::P + k;
                const T v = q[t] * factor<T>(wv[t], has_w, coef);
                if (j & 1) s->y = v;
                else s->x = v;
            }
        }
        __syncthreads();
        ct_slab_forward<T, T2, N>(S, Wy, Wz);
        T2* c = (o == 0 ? it.c0 : it.c1) + coff;
        for (int f = threadIdx.x; f < G::NSPEC; f += BD) c[f] = S[(f / G::NZC) * G::P + f % G::NZC];
    }
}

// ------------------------------------------------------------------ pencil kernel, N^3
template <int N>
struct PencilShape {
    static constexpr int NZC = N / 2 + 1;
    // kz tile: N * tiles >= 128 workgroups per item, tiles of at least 8
    static constexpr int tiles() {
        int C = NZC, t = 1;
        while (N * t < 128 && C > 8) {
            C = (C + 1) / 2;
            t = (NZC + C - 1) / C;
        }
        return t;
    }
    static constexpr int TILES = tiles();
    static constexpr int C = (NZC + TILES - 1) / TILES;
    static constexpr int P = C;
    static constexpr int BD = ((N * C + 7) / 8 + 63) / 64 * 64 > 1024 ? 1024 : ((N * C + 7) / 8 + 63) / 64 * 64;
    static constexpr size_t lds_elems = (size_t)N * P + N;
};

template <typename T, int N>
__global__ void __launch_bounds__(PencilShape<N>::BD) xpass_kernel_ct(XBatch<T> batch, const typename vec2<T>::type* twx,
                                                                     const T* bx, const T* by, const T* bz) {
    using T2 = typename vec2<T>::type;
    using G = PencilShape<N>;
    constexpr int BD = G::BD, C = G::C, NZC = G::NZC;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T2* S = reinterpret_cast<T2*>(smem_raw);
    T2* Wx = S + N * G::P;
    for (int m = threadIdx.x; m < N; m += BD) Wx[m] = twx[m];
    const XItem<T>& it = batch.it[blockIdx.y];
    const int j = blockIdx.x / G::TILES;
    const int tile = blockIdx.x - j * G::TILES;
    const int k0 = tile * C;
    const int cw = min(C, NZC - k0);
    constexpr size_t ROW = (size_t)N * NZC;
    T2* c = it.c + (size_t)j * NZC + k0;
    constexpr int TOTAL = N * C;
    // the last tile may be narrower: its unused columns are loaded as zeros
    for (int f = threadIdx.x; f < TOTAL; f += BD) {
        const int i = f / C, k = f - i * C;
        S[i * G::P + k] = k < cw ? c[i * ROW + k] : T2{0, 0};
    }
    __syncthreads();
    if (it.table != -2) ct_fft<-1, T, T2, N, C, 1, G::P, BD>(S, Wx);
    if (it.table >= 0) {
        const T* tx = bx + (size_t)it.table * N;
        const T byj = by[(size_t)it.table * N + j];
        const T* tz = bz + (size_t)it.table * NZC + k0;
        for (int f = threadIdx.x; f < TOTAL; f += BD) {
            const int i = f / C, k = f - i * C;
            if (k < cw) {
                const T b = tx[i] * byj * tz[k];
                T2 v = S[i * G::P + k];
                S[i * G::P + k] = T2{v.x * b, v.y * b};
            }
        }
        __syncthreads();
    }
    if (it.table != -1) ct_fft<+1, T, T2, N, C, 1, G::P, BD>(S, Wx);
    for (int f = threadIdx.x; f < TOTAL; f += BD) {
        const int i = f / C, k = f - i * C;
        if (k < cw) c[i * ROW + k] = S[i * G::P + k];
    }
}

}  // namespace fts
