// Wave-local slab and pencil kernels for cubic grids N = 32, 48, 64, 96, 128.
//
// Every 1D line transform belongs to one wavefront for all its Stockham stages:
// a stage's butterflies take their inputs from registers, and the exchange between
// stages goes through the wave's own LDS lines with no workgroup barrier (LDS
// operations of one wavefront complete in issue order; wave_sync keeps the
// compiler from moving them). Registers carry data across phases wherever the
// output pattern of one transform is the input pattern of the next: a stage of
// radix R reads positions j + r L/R of its line, and the last stage of a plan
// writes exactly that pattern for its own radix, so a transform followed by the
// reversed plan needs no exchange at the junction.
//
// Slab pass (one x-slab per workgroup): 3 workgroup barriers per spectrum (twiddles
// and plane reuse, y columns -> z rows, z rows -> y columns). Pencil pass (one
// wavefront per y row and 8 kz columns): no barrier at all.
#pragma once

#include "fts_fft_ct.hpp"

namespace fts {

// Stores of results that the next pass reads: plain, or streaming (nontemporal) when
// FTS_STREAM_STORES is 1.
#ifndef FTS_STREAM_STORES
#define FTS_STREAM_STORES 0
#endif
template <typename T>
__device__ __forceinline__ void put(T* p, T v) {
    if constexpr (FTS_STREAM_STORES) __builtin_nontemporal_store(v, p);
    else *p = v;
}
template <typename T2>
__device__ __forceinline__ void put2(T2* p, T2 v) {
    if constexpr (FTS_STREAM_STORES) {
        __builtin_nontemporal_store(v.x, &p->x);
        __builtin_nontemporal_store(v.y, &p->y);
    } else {
        *p = v;
    }
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// radix plan of a length-L transform, optionally reversed
template <int L, bool REV>
struct WPlan {
    static constexpr CPlan P = make_cplan(L);
    static constexpr int n = P.n;
    static constexpr int r(int s) { return REV ? P.r[P.n - 1 - s] : P.r[s]; }
    static constexpr int ns(int s) { return s == 0 ? 1 : ns(s - 1) * r(s - 1); }
};

// Butterfly g of a stage with radix R over NLW lines of a wavefront: lane g % 64,
// register slot g / 64; line-fast order (LF) puts consecutive lanes on consecutive
// lines, otherwise on consecutive butterflies of one line.
template <int L, int R, int NLW, bool LF>
struct WMap {
    static constexpr int B = L / R;
    static constexpr int TOTAL = NLW * B;
    static constexpr int NT = (TOTAL + 63) / 64;
    static __device__ __forceinline__ int line(int g) { return LF ? g % NLW : g / B; }
    static __device__ __forceinline__ int idx(int g) { return LF ? g / NLW : g % B; }
    static __device__ __forceinline__ bool ok(int g, int nvalid) {
        return (TOTAL % 64 == 0 || g < TOTAL) && line(g) < nvalid;
    }
};

template <int L, bool REV, int S, int NLW, bool LF>
using WStageMap = WMap<L, WPlan<L, REV>::r(S), NLW, LF>;

// registers of one stage's inputs: [slots][radix]
template <typename T2, int L, bool REV, int S, int NLW, bool LF>
struct WRegs {
    using type = T2[WStageMap<L, REV, S, NLW, LF>::NT][WPlan<L, REV>::r(S)];
};

// Position of element i of a line in the exchange between two stages: with SK one pad
// slot after every 8 elements, so that the stride-R stores of a stage's butterflies
// (positions R j + r) fall on different LDS banks; the line then spans L + (L-1)/8 slots.
template <bool SK>
__device__ __forceinline__ int wl_slot(int i) {
    return SK ? i + (i >> 3) : i;
}

// Stages S.. of the plan over this wavefront's lines (line l at X + l LP, element i
// at stride ES; skewed exchange positions with SK). v holds stage S's inputs; `done`
// receives the last stage's outputs in registers (positions j + r L/R_last).
template <int DIR, typename T, typename T2, int L, bool REV, int NLW, bool LF, int LP, int ES, bool SK, int S,
          typename F>
__device__ __forceinline__ void wl_stages(typename WRegs<T2, L, REV, S, NLW, LF>::type& v, T2* X, const T2* W,
                                          int lane, int nvalid, F&& done) {
    using PL = WPlan<L, REV>;
    constexpr int R = PL::r(S);
    constexpr int NS = PL::ns(S);
    constexpr int TWS = L / (NS * R);
    using M = WMap<L, R, NLW, LF>;
#pragma unroll
    for (int t = 0; t < M::NT; t++) {
        const int g = lane + 64 * t;
        if (M::ok(g, nvalid)) {
            if constexpr (NS > 1) {
                const int k = M::idx(g) % NS;
#pragma unroll
                for (int r = 1; r < R; r++) v[t][r] = cmul(v[t][r], root<DIR>(W, (r * k * TWS) % L));
            }
            Dft<R, DIR, T, T2>::run(v[t], W, L);
        }
    }
    if constexpr (S + 1 == PL::n) {
        done(v);
    } else {
        wave_sync();
#pragma unroll
        for (int t = 0; t < M::NT; t++) {
            const int g = lane + 64 * t;
            if (M::ok(g, nvalid)) {
                const int j = M::idx(g);
                const int d = (j / NS) * NS * R + j % NS;
                T2* base = X + M::line(g) * LP;
#pragma unroll
                for (int r = 0; r < R; r++) base[wl_slot<SK>(d + r * NS) * ES] = v[t][r];
            }
        }
        wave_sync();
        using M2 = WStageMap<L, REV, S + 1, NLW, LF>;
        constexpr int R2 = PL::r(S + 1);
        T2 u[M2::NT][R2];
#pragma unroll
        for (int t = 0; t < M2::NT; t++) {
            const int g = lane + 64 * t;
            if (M2::ok(g, nvalid)) {
                const T2* base = X + M2::line(g) * LP;
                const int j = M2::idx(g);
#pragma unroll
                for (int r = 0; r < R2; r++) u[t][r] = base[wl_slot<SK>(j + r * M2::B) * ES];
            }
        }
        wl_stages<DIR, T, T2, L, REV, NLW, LF, LP, ES, SK, S + 1>(u, X, W, lane, nvalid, done);
    }
}

// ------------------------------------------------------------------ slab pass
// Launch shapes of the wave-local kernels. Each (N, precision) picks its slab workgroup
// size and pencil width from the candidates below when a grid is created, by timing
// them (Grid::tune_fft_schedule); every candidate runs the same per-line transforms,
// so results agree to rounding (bit-identical at 32^3 and 64^3). The defaults are used
// when tuning is off (FTS_FFT_TUNE=0).
//   64^3 slab default: 5 wavefronts own 7 y columns / 7 packed z rows each (88% of the
//   lanes busy in both phases).
constexpr int wslab_default_bd(int N) {
    return N <= 32 ? 256 : (N <= 48 ? 384 : (N <= 64 ? 320 : 1024));
}
constexpr int WPENCIL_DEFAULT_NL = 8;
// default kz columns per pencil wavefront for an N^3 grid and complex element size: 4 in
// FP64 from 96^3 (more wavefronts; the XCD tile order keeps a line's tiles on one L2),
// measured 4-6% faster per step than 8 there and slower at 64^3
constexpr int wpencil_default_nl(int N, int elem_bytes) {
    return elem_bytes == 16 && N >= 96 ? 4 : WPENCIL_DEFAULT_NL;
}

// Plane row pitch (complex elements) of the slab pass. A y column exchange puts the
// NLY columns of a wavefront on consecutive lanes and the next butterfly on the next
// row, so a pitch of NLY (mod 8) keeps 8 consecutive lanes' 16-byte slots on distinct
// banks; the pitch also holds a skewed packed z line in two rows (2P >= N + (N-1)/8).
// Chosen with a bank model of every LDS access of the pass (48^3: one more step of 8,
// better for the 64-bank case of that model).
// FTS_SLAB_PAD=0 (a tuning define) keeps the unpadded pitch N/2+1 and unskewed z lines.
#ifndef FTS_SLAB_PAD
#define FTS_SLAB_PAD 1
#endif
constexpr bool WSLAB_SKEW = FTS_SLAB_PAD != 0;
constexpr int wslab_pitch(int N, int NLY) {
    if (!WSLAB_SKEW) return N / 2 + 1;
    const int lo0 = (N + (N - 1) / 8 + 1) / 2;
    const int lo = lo0 > N / 2 + 1 ? lo0 : N / 2 + 1;
    const int P = lo + ((NLY - lo) % 8 + 8) % 8;
    return (N == 48 && NLY == 5) ? P + 8 : P;
}

// Row pitch (complex elements T2) of the half spectra that the wave-local passes exchange
// through HBM: nz/2+1 rounded up to a whole number of 128-byte lines (8 columns in FP64,
// 16 in FP32), so that every spectrum row, and with it every pencil tile of that many
// columns, starts on a line. A tile that straddles two lines, or shares one with a
// neighbouring tile on another XCD (its own L2), fetches lines twice. The pad columns are
// never read or written. FTS_SPEC_PAD=0 (a tuning define) keeps nz/2+1.
#ifndef FTS_SPEC_PAD
#define FTS_SPEC_PAD 1
#endif
template <typename T2>
constexpr int wspec_pitch(int N) {
    constexpr int L = 128 / (int)sizeof(T2);
    return FTS_SPEC_PAD ? (N / 2 + 1 + L - 1) / L * L : N / 2 + 1;
}

template <int N, int BD_ = wslab_default_bd(N)>
struct WSlab {
    static constexpr int NZC = N / 2 + 1;
    static constexpr int NP = N / 2;
    static constexpr int BD = BD_;
    static constexpr int NW = BD / 64;
    static constexpr int NLY = (NZC + NW - 1) / NW;   // y columns per wavefront
    static constexpr int NLZ = (NP + NW - 1) / NW;    // packed z rows per wavefront
    static constexpr int P = wslab_pitch(N, NLY);     // plane row pitch; packed z line p spans rows 2p, 2p+1
    static constexpr size_t lds_elems = (size_t)N * P + 2 * N;
    static_assert(BD % 64 == 0 && BD <= 1024, "whole wavefronts, at most 1024 threads");
    static_assert(NW * NLY >= NZC && NW * NLZ >= NP, "the wavefronts cover every y column and packed z row");
    static_assert(P >= NZC && (!WSLAB_SKEW || 2 * P >= N + (N - 1) / 8),
                  "a row holds a half spectrum, two rows a skewed z line");
    static_assert(lds_elems * 16 <= 160 * 1024, "the FP64 plane and twiddles fit the LDS");
    using YI = WPlan<N, false>;
    using YF = WPlan<N, false>;
    using ZI = WPlan<N, false>;
    using ZF = WPlan<N, true>;   // first radix = last radix of ZI: takes ZI's output registers
    static constexpr int RZ = ZI::r(ZI::n - 1);
    using MZ = WMap<N, RZ, NLZ, false>;   // real values of a wavefront's packed rows
};

template <typename T, int MODE, int N, int BD>
__global__ void __launch_bounds__(BD) slab_kernel_wl(SlabBatch<T> batch, const typename vec2<T>::type* twy,
                                                     const typename vec2<T>::type* twz, double inv_m_d) {
    using T2 = typename vec2<T>::type;
    using G = WSlab<N, BD>;
    using MZ = typename G::MZ;
    constexpr int RZ = G::RZ, NZC = G::NZC, P = G::P;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T2* S = reinterpret_cast<T2*>(smem_raw);
    T2* Wy = S + N * P;
    T2* Wz = Wy + N;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const SlabItem<T>& it = batch.it[blockIdx.y];
    const size_t roff = (size_t)blockIdx.x * N * N;
    constexpr int SP = wspec_pitch<T2>(N);            // spectrum row pitch in HBM
    const size_t coff = (size_t)blockIdx.x * N * SP;  // one x-slab of the spectrum
    const T inv_m = (T)inv_m_d;
    const bool has_w = it.w != nullptr;
    const int col0 = wave * G::NLY;
    const int ncol = max(0, min(G::NLY, NZC - col0));
    const int p0 = wave * G::NLZ;
    const int nrow = max(0, min(G::NLZ, G::NP - p0));

    // w at this wavefront's real positions: rows 2p (e) and 2p + 1 (o), column k.
    // Up to N = 64 it is loaded once and stays in registers for the whole kernel. A
    // 96^3 or 128^3 real plane takes 24-32 VGPRs per thread at 1024 threads, and
    // holding w beside the propagator's own plane and a transform's working set
    // spilled to scratch; there each phase that applies a factor reloads w (an L2
    // hit after the first read), so w is live only while the phase runs.
    constexpr bool W_RESIDENT = N <= 64;
    // The y phase of a wavefront runs its NLY columns in YNCH chunks of YCH columns, so
    // that a chunk's first-stage butterflies fill at most one register slot per lane:
    // at 128^3 with 1024 threads all 5 columns at once take two slots (80 butterflies
    // for 64 lanes) and double the phase's registers. One chunk up to N = 64.
    constexpr int YB0 = N / G::YI::r(0);
    constexpr int YNCH = W_RESIDENT ? 1 : (G::NLY * YB0 + 63) / 64;
    constexpr int YCH = (G::NLY + YNCH - 1) / YNCH;
    auto load_w = [&](T (&we)[W_RESIDENT ? MZ::NT : 1][RZ], T (&wo)[W_RESIDENT ? MZ::NT : 1][RZ]) {
#pragma unroll
        for (int t = 0; t < MZ::NT; t++) {
            const int g = lane + 64 * t;
            const bool ok = MZ::ok(g, nrow);
            const int p = p0 + MZ::line(g);
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int k = MZ::idx(g) + r * MZ::B;
                we[t][r] = (has_w && ok) ? it.w[roff + (size_t)(2 * p) * N + k] : (T)0;
                wo[t][r] = (has_w && ok) ? it.w[roff + (size_t)(2 * p + 1) * N + k] : (T)0;
            }
        }
    };
    T we_r[W_RESIDENT ? MZ::NT : 1][RZ], wo_r[W_RESIDENT ? MZ::NT : 1][RZ];
    if constexpr (W_RESIDENT) load_w(we_r, wo_r);
    for (int m = threadIdx.x; m < N; m += G::BD) {
        Wy[m] = twy[m];
        Wz[m] = twz[m];
    }

    // spectrum -> real rows times exp(-w coef) / M, in registers
    auto inverse = [&](const T2* spec, T coef, T (&qe)[MZ::NT][RZ], T (&qo)[MZ::NT][RZ], auto comb) {
        using MA = WStageMap<N, false, 0, YCH, true>;
#pragma unroll
        for (int h = 0; h < YNCH; h++) {
            const int cb = col0 + h * YCH;
            const int nc = max(0, min(YCH, ncol - h * YCH));
            T2 v[MA::NT][G::YI::r(0)];
#pragma unroll
            for (int t = 0; t < MA::NT; t++) {
                const int g = lane + 64 * t;
                if (MA::ok(g, nc)) {
                    const int c = cb + MA::line(g), j = MA::idx(g);
#pragma unroll
                    for (int r = 0; r < G::YI::r(0); r++) v[t][r] = spec[(j + r * MA::B) * SP + c];
                }
            }
            if (h == 0) __syncthreads();   // twiddles stored; the plane's previous users are done
            wl_stages<+1, T, T2, N, false, YCH, true, 1, P, false, 0>(v, S + cb, Wy, lane, nc, [&](auto& o) {
                using MO = WStageMap<N, false, G::YI::n - 1, YCH, true>;
                wave_sync();
#pragma unroll
                for (int t = 0; t < MO::NT; t++) {
                    const int g = lane + 64 * t;
                    if (MO::ok(g, nc)) {
                        const int c = cb + MO::line(g), j = MO::idx(g);
#pragma unroll
                        for (int r = 0; r < G::YI::r(G::YI::n - 1); r++) S[(j + r * MO::B) * P + c] = o[t][r];
                    }
                }
            });
        }
        __syncthreads();
        // packed rows: Z[k] = X_{2p}[k] + i X_{2p+1}[k] from the half spectra (Hermitian fill)
        using MB = WStageMap<N, false, 0, G::NLZ, false>;
        T2 u[MB::NT][G::ZI::r(0)];
#pragma unroll
        for (int t = 0; t < MB::NT; t++) {
            const int g = lane + 64 * t;
            if (MB::ok(g, nrow)) {
                const int p = p0 + MB::line(g), j = MB::idx(g);
#pragma unroll
                for (int r = 0; r < G::ZI::r(0); r++) {
                    const int k = j + r * MB::B;
                    const bool upper = k >= NZC;
                    const int kk = upper ? N - k : k;
                    T2 a = S[2 * p * P + kk];
                    T2 b = S[(2 * p + 1) * P + kk];
                    if (kk == 0 || 2 * kk == N) {
                        a.y = 0;
                        b.y = 0;
                    }
                    if (upper) {
                        a.y = -a.y;
                        b.y = -b.y;
                    }
                    u[t][r] = T2{a.x - b.y, a.y + b.x};
                }
            }
        }
        // q = o exp(-w coef) / M; with COMB the Richardson combination (4 q - q_prev) / 3
        // against the q_prev already in qe/qo. Without resident w, each register slot
        // loads its w just before use.
        auto finish = [&](auto& o) {
#pragma unroll
            for (int t = 0; t < MZ::NT; t++) {
                const int g = lane + 64 * t;
                if (MZ::ok(g, nrow)) {
                    T we[RZ], wo[RZ];
                    if constexpr (W_RESIDENT) {
#pragma unroll
                        for (int r = 0; r < RZ; r++) {
                            we[r] = we_r[t][r];
                            wo[r] = wo_r[t][r];
                        }
                    } else {
                        const int p = p0 + MZ::line(g);
#pragma unroll
                        for (int r = 0; r < RZ; r++) {
                            const int k = MZ::idx(g) + r * MZ::B;
                            we[r] = has_w ? it.w[roff + (size_t)(2 * p) * N + k] : (T)0;
                            wo[r] = has_w ? it.w[roff + (size_t)(2 * p + 1) * N + k] : (T)0;
                        }
                    }
#pragma unroll
                    for (int r = 0; r < RZ; r++) {
                        const T ve = o[t][r].x * factor<T>(we[r], has_w, coef) * inv_m;
                        const T vo = o[t][r].y * factor<T>(wo[r], has_w, coef) * inv_m;
                        const size_t ie = roff + (size_t)(2 * (p0 + MZ::line(g))) * N + MZ::idx(g) + r * MZ::B;
                        if constexpr (decltype(comb)::value == 1) {
                            qe[t][r] = ((T)4.0 * ve - qe[t][r]) / (T)3.0;
                            qo[t][r] = ((T)4.0 * vo - qo[t][r]) / (T)3.0;
                        } else if constexpr (decltype(comb)::value == 2) {
                            // q1 of the Richardson step parked in the output slice
                            it.qout[ie] = ve;
                            it.qout[ie + N] = vo;
                        } else if constexpr (decltype(comb)::value == 3) {
                            qe[t][r] = ((T)4.0 * ve - it.qout[ie]) / (T)3.0;
                            qo[t][r] = ((T)4.0 * vo - it.qout[ie + N]) / (T)3.0;
                        } else {
                            qe[t][r] = ve;
                            qo[t][r] = vo;
                        }
                    }
                }
            }
        };
        wl_stages<+1, T, T2, N, false, G::NLZ, false, 2 * P, 1, WSLAB_SKEW, 0>(u, S + 2 * p0 * P, Wz, lane, nrow, finish);
    };

    // real rows times exp(-w coef) -> spectrum
    auto forward = [&](const T (&qe)[MZ::NT][RZ], const T (&qo)[MZ::NT][RZ], T coef, T2* spec) {
        T2 u[MZ::NT][RZ];
#pragma unroll
        for (int t = 0; t < MZ::NT; t++) {
            const int g = lane + 64 * t;
            if (MZ::ok(g, nrow)) {
                T we[RZ], wo[RZ];
                if constexpr (W_RESIDENT) {
#pragma unroll
                    for (int r = 0; r < RZ; r++) {
                        we[r] = we_r[t][r];
                        wo[r] = wo_r[t][r];
                    }
                } else {
                    const int p = p0 + MZ::line(g);
#pragma unroll
                    for (int r = 0; r < RZ; r++) {
                        const int k = MZ::idx(g) + r * MZ::B;
                        we[r] = has_w ? it.w[roff + (size_t)(2 * p) * N + k] : (T)0;
                        wo[r] = has_w ? it.w[roff + (size_t)(2 * p + 1) * N + k] : (T)0;
                    }
                }
#pragma unroll
                for (int r = 0; r < RZ; r++)
                    u[t][r] = T2{qe[t][r] * factor<T>(we[r], has_w, coef), qo[t][r] * factor<T>(wo[r], has_w, coef)};
            }
        }
        wl_stages<-1, T, T2, N, true, G::NLZ, false, 2 * P, 1, WSLAB_SKEW, 0>(u, S + 2 * p0 * P, Wz, lane, nrow, [&](auto& o) {
            using MO = WStageMap<N, true, G::ZF::n - 1, G::NLZ, false>;
            wave_sync();
#pragma unroll
            for (int t = 0; t < MO::NT; t++) {
                const int g = lane + 64 * t;
                if (MO::ok(g, nrow)) {
                    const int p = p0 + MO::line(g), j = MO::idx(g);
#pragma unroll
                    for (int r = 0; r < G::ZF::r(G::ZF::n - 1); r++) S[2 * p * P + j + r * MO::B] = o[t][r];
                }
            }
        });
        wave_sync();
        // split Z into the half spectra of rows 2p and 2p + 1
        constexpr int TOT = G::NLZ * NZC;
        constexpr int NT = (TOT + 63) / 64;
        T2 va[NT], vb[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int f = lane + 64 * t;
            const int l = f / NZC, k = f - l * NZC;
            if ((TOT % 64 == 0 || f < TOT) && l < nrow) {
                const T2* Z = S + 2 * (p0 + l) * P;
                const T2 A = Z[k];
                const int km = k == 0 ? 0 : N - k;
                const T2 B = T2{Z[km].x, -Z[km].y};
                va[t] = T2{(T)0.5 * (A.x + B.x), (T)0.5 * (A.y + B.y)};
                vb[t] = T2{(T)0.5 * (A.y - B.y), (T)-0.5 * (A.x - B.x)};
            }
        }
        wave_sync();
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int f = lane + 64 * t;
            const int l = f / NZC, k = f - l * NZC;
            if ((TOT % 64 == 0 || f < TOT) && l < nrow) {
                S[2 * (p0 + l) * P + k] = va[t];
                S[(2 * (p0 + l) + 1) * P + k] = vb[t];
            }
        }
        __syncthreads();
        using MD = WStageMap<N, false, 0, YCH, true>;
#pragma unroll
        for (int h = 0; h < YNCH; h++) {
            const int cb = col0 + h * YCH;
            const int nc = max(0, min(YCH, ncol - h * YCH));
            T2 v[MD::NT][G::YF::r(0)];
#pragma unroll
            for (int t = 0; t < MD::NT; t++) {
                const int g = lane + 64 * t;
                if (MD::ok(g, nc)) {
                    const int c = cb + MD::line(g), j = MD::idx(g);
#pragma unroll
                    for (int r = 0; r < G::YF::r(0); r++) v[t][r] = S[(j + r * MD::B) * P + c];
                }
            }
            wl_stages<-1, T, T2, N, false, YCH, true, 1, P, false, 0>(v, S + cb, Wy, lane, nc, [&](auto& o) {
                using MO = WStageMap<N, false, G::YF::n - 1, YCH, true>;
#pragma unroll
                for (int t = 0; t < MO::NT; t++) {
                    const int g = lane + 64 * t;
                    if (MO::ok(g, nc)) {
                        const int c = cb + MO::line(g), j = MO::idx(g);
#pragma unroll
                        for (int r = 0; r < G::YF::r(G::YF::n - 1); r++)
                            put2(&spec[(j + r * MO::B) * SP + c], o[t][r]);
                    }
                }
            });
        }
    };

    T qe[MZ::NT][RZ], qo[MZ::NT][RZ];
    bool do_fwd = true;
    if constexpr (MODE == SLAB_FWD) {
#pragma unroll
        for (int t = 0; t < MZ::NT; t++) {
            const int g = lane + 64 * t;
            const bool ok = MZ::ok(g, nrow);
            const int p = p0 + MZ::line(g);
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int k = MZ::idx(g) + r * MZ::B;
                qe[t][r] = ok ? it.qin[roff + (size_t)(2 * p) * N + k] : (T)0;
                qo[t][r] = ok ? it.qin[roff + (size_t)(2 * p + 1) * N + k] : (T)0;
            }
        }
        __syncthreads();   // twiddles
    } else {
        // Richardson step: (4 q2 - q1) / 3 with q2 from the second spectrum. Without
        // resident w, q1 waits in the output slice (always set for INV2 items at these
        // sizes; the host checks it) instead of in registers through the second inverse.
        using Comb = std::integral_constant<int, W_RESIDENT ? 1 : 3>;
        using First = std::integral_constant<int, (MODE == SLAB_INV2 && !W_RESIDENT) ? 2 : 0>;
        inverse(it.c0 + coff, it.post0, qe, qo, First{});
        if constexpr (MODE == SLAB_INV2) inverse(it.c1 + coff, it.post1, qe, qo, Comb{});
#pragma unroll
        for (int t = 0; t < MZ::NT; t++) {
            const int g = lane + 64 * t;
            if (MZ::ok(g, nrow)) {
                const int p = p0 + MZ::line(g);
#pragma unroll
                for (int r = 0; r < RZ; r++) {
                    const int k = MZ::idx(g) + r * MZ::B;
                    const size_t ie = roff + (size_t)(2 * p) * N + k, io = ie + N;
                    if (it.mask) {
                        qe[t][r] *= it.mask[ie];
                        qo[t][r] *= it.mask[io];
                    }
                    if (it.qout) {
                        put(&it.qout[ie], qe[t][r]);
                        put(&it.qout[io], qo[t][r]);
                    }
                }
            }
        }
        do_fwd = it.forward != 0;
    }
    if (!do_fwd) return;
    const int n_out = (MODE == SLAB_FWD) ? it.n_out : (MODE == SLAB_INV2 ? 2 : 1);
    forward(qe, qo, it.pre0, it.c0 + coff);
    if (n_out < 2) return;
    __syncthreads();   // the plane of the first spectrum is fully read
    if constexpr (!W_RESIDENT) {
        // the second forward transform starts from the real plane this thread read
        // (FWD: qin) or wrote (INV2: qout, always set for INV2 items at these sizes;
        // the host checks it), so the plane is not held in registers through the first
#pragma unroll
        for (int t = 0; t < MZ::NT; t++) {
            const int g = lane + 64 * t;
            const bool ok = MZ::ok(g, nrow);
            const int p = p0 + MZ::line(g);
            const T* src = MODE == SLAB_FWD ? it.qin : it.qout;
#pragma unroll
            for (int r = 0; r < RZ; r++) {
                const int k = MZ::idx(g) + r * MZ::B;
                qe[t][r] = ok ? src[roff + (size_t)(2 * p) * N + k] : (T)0;
                qo[t][r] = ok ? src[roff + (size_t)(2 * p + 1) * N + k] : (T)0;
            }
        }
    }
    forward(qe, qo, it.pre1, it.c1 + coff);
}

// ------------------------------------------------------------------ pencil pass
template <int N, int NL_ = WPENCIL_DEFAULT_NL>
struct WPencil {
    static constexpr int NZC = N / 2 + 1;
    static constexpr int NL = NL_;   // kz columns per wavefront
    static constexpr int TILES = (NZC + NL - 1) / NL;
    static constexpr int BD = 64;
    static constexpr size_t lds_elems = (size_t)N * NL + N;
    using F = WPlan<N, false>;
    using I = WPlan<N, true>;   // first radix = last radix of F
};

// Pencil tiles narrower than a 128-byte line share that line of a spectrum row (rows start
// on a line) with SH - 1 neighbours. Workgroups go to the 8 XCDs round-robin by linear id,
// and each XCD has its own L2, so neighbouring tiles would each fetch the line. The grid
// then pads a row's tiles to a whole number of lines, and xcd_tile orders the tiles so that
// the SH tiles of one line get ids 8 apart: one XCD, one fetch.
template <typename T2, int NL>
constexpr int wpencil_share() {
    return 128 / (NL * (int)sizeof(T2)) > 1 ? 128 / (NL * (int)sizeof(T2)) : 1;
}
template <typename T2, int N, int NL>
constexpr int wpencil_tiles() {
    constexpr int SH = wpencil_share<T2, NL>();
    return ((N / 2 + 1 + NL - 1) / NL + SH - 1) / SH * SH;
}
// tile of block b: b = 8 SH g + 8 s + k (k < 8, s < SH) takes tile SH (8 g + k) + s; a
// bijection when the grid's x extent is a multiple of 8 SH (N x a multiple of SH, N % 8 == 0)
template <int SH>
__device__ __forceinline__ int xcd_tile(int b) {
    if constexpr (SH == 1) {
        return b;
    } else {
        const int g = b / (8 * SH), r = b - g * 8 * SH;
        return SH * (8 * g + (r & 7)) + (r >> 3);
    }
}

template <typename T, int N, int NLW>
__global__ void __launch_bounds__(64) xpass_kernel_wl(XBatch<T> batch, const typename vec2<T>::type* twx,
                                                      const T* bx, const T* by, const T* bz) {
    using T2 = typename vec2<T>::type;
    using G = WPencil<N, NLW>;
    using F = typename G::F;
    using I = typename G::I;
    constexpr int NL = G::NL, NZC = G::NZC;
    constexpr int TP = wpencil_tiles<T2, N, NL>();
    static_assert(N % 8 == 0, "the XCD tile order needs a grid extent of a multiple of 8 lines' tiles");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T2* S = reinterpret_cast<T2*>(smem_raw);   // line l, element i at S[i NL + l]
    T2* Wx = S + N * NL;
    const int lane = threadIdx.x;
    const XItem<T>& it = batch.it[blockIdx.y];
    const int tb = xcd_tile<wpencil_share<T2, NL>()>(blockIdx.x);
    const int j = tb / TP;
    const int k0 = (tb - j * TP) * NL;
    if (k0 >= NZC) return;   // a pad tile: the whole wavefront leaves
    const int nl = min(NL, NZC - k0);
    constexpr int SP = wspec_pitch<T2>(N);
    constexpr size_t ROW = (size_t)N * SP;
    T2* c = it.c + (size_t)j * SP + k0;
    const int table = it.table;
    for (int m = lane; m < N; m += 64) Wx[m] = twx[m];

    auto store = [&](auto& o) {
        using MO = WStageMap<N, true, I::n - 1, NL, true>;
#pragma unroll
        for (int t = 0; t < MO::NT; t++) {
            const int g = lane + 64 * t;
            if (MO::ok(g, nl)) {
                const int l = MO::line(g), i0 = MO::idx(g);
#pragma unroll
                for (int r = 0; r < I::r(I::n - 1); r++) put2(&c[(i0 + r * MO::B) * ROW + l], o[t][r]);
            }
        }
    };
    if (table == -2) {
        using M0 = WStageMap<N, true, 0, NL, true>;
        T2 v[M0::NT][I::r(0)];
#pragma unroll
        for (int t = 0; t < M0::NT; t++) {
            const int g = lane + 64 * t;
            if (M0::ok(g, nl)) {
                const int l = M0::line(g), i0 = M0::idx(g);
#pragma unroll
                for (int r = 0; r < I::r(0); r++) v[t][r] = c[(i0 + r * M0::B) * ROW + l];
            }
        }
        wave_sync();
        wl_stages<+1, T, T2, N, true, NL, true, 1, NL, false, 0>(v, S, Wx, lane, nl, store);
        return;
    }
    using M0 = WStageMap<N, false, 0, NL, true>;
    T2 v[M0::NT][F::r(0)];
#pragma unroll
    for (int t = 0; t < M0::NT; t++) {
        const int g = lane + 64 * t;
        if (M0::ok(g, nl)) {
            const int l = M0::line(g), i0 = M0::idx(g);
#pragma unroll
            for (int r = 0; r < F::r(0); r++) v[t][r] = c[(i0 + r * M0::B) * ROW + l];
        }
    }
    wave_sync();
    wl_stages<-1, T, T2, N, false, NL, true, 1, NL, false, 0>(v, S, Wx, lane, nl, [&](auto& o) {
        using MF = WStageMap<N, false, F::n - 1, NL, true>;
        constexpr int RF = F::r(F::n - 1);
        if (table == -1) {
#pragma unroll
            for (int t = 0; t < MF::NT; t++) {
                const int g = lane + 64 * t;
                if (MF::ok(g, nl)) {
                    const int l = MF::line(g), i0 = MF::idx(g);
#pragma unroll
                    for (int r = 0; r < RF; r++) put2(&c[(i0 + r * MF::B) * ROW + l], o[t][r]);
                }
            }
            return;
        }
        const T* tx = bx + (size_t)table * N;
        const T byj = by[(size_t)table * N + j];
        const T* tz = bz + (size_t)table * NZC + k0;
#pragma unroll
        for (int t = 0; t < MF::NT; t++) {
            const int g = lane + 64 * t;
            if (MF::ok(g, nl)) {
                const int l = MF::line(g), i0 = MF::idx(g);
                const T bl = byj * tz[l];
#pragma unroll
                for (int r = 0; r < RF; r++) {
                    const T b = tx[i0 + r * MF::B] * bl;
                    o[t][r] = T2{o[t][r].x * b, o[t][r].y * b};
                }
            }
        }
        wl_stages<+1, T, T2, N, true, NL, true, 1, NL, false, 0>(o, S, Wx, lane, nl, store);
    });
}

}  // namespace fts
