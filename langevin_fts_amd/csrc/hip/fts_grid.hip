// Grid (ComputationBox + FFT plan), separable Boltzmann tables and the
// deterministic reductions (replacing cub::DeviceReduce::Sum, CudaComputationBox.cu:56-118).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>

#include "fts_internal.hpp"
#include "fts_reduce.hpp"

namespace fts {

// ------------------------------------------------------------------ kernel timer
KernelTimer::~KernelTimer() {
    for (auto& p : pending_) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : pool_) (void)hipEventDestroy(e);
}

void KernelTimer::enable(int sample_every) {
    drain(pending_.size());
    for (auto& s : stats_) s = fts_kernel_stats{};
    for (auto& c : count_) c = 0;
    every_ = sample_every > 0 ? sample_every : 0;
}

hipEvent_t KernelTimer::take() {
    if (pool_.empty()) {
        // timing-only events: no system-scope fence (no L2 writeback/invalidate
        // between the timed kernels)
        hipEvent_t e;
        FTS_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        return e;
    }
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
}

hipEvent_t KernelTimer::begin(hipStream_t st, int family) {
    if (every_ == 0 || count_[family]++ % every_ != 0) return nullptr;
    hipEvent_t a = take();
    FTS_HIP_CHECK(hipEventRecord(a, st));
    return a;
}

void KernelTimer::end(hipStream_t st, hipEvent_t a, int family, double bytes) {
    hipEvent_t b = take();
    FTS_HIP_CHECK(hipEventRecord(b, st));
    pending_.push_back({a, b, family, bytes});
    // the older half finished long ago: reading it never waits on the GPU's queue
    if (pending_.size() >= 8192) drain(pending_.size() / 2);
}

void KernelTimer::drain(size_t n) {
    for (size_t i = 0; i < n; i++) {
        Pending& p = pending_[i];
        FTS_HIP_CHECK(hipEventSynchronize(p.b));
        float ms = 0.f;
        FTS_HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
        stats_[p.family].milliseconds += ms;
        stats_[p.family].bytes += p.bytes;
        stats_[p.family].launches += 1;
        pool_.push_back(p.a);
        pool_.push_back(p.b);
    }
    pending_.erase(pending_.begin(), pending_.begin() + n);
}

fts_kernel_stats KernelTimer::get(int family) {
    drain(pending_.size());
    return stats_[family];
}

namespace {

// ~8 elements per thread, whole waves, at most 1024 threads
int choose_threads(int work) {
    int bd = ceil_div(ceil_div(work, 8), 64) * 64;
    return std::max(64, std::min(1024, bd));
}

template <typename T>
void upload_twiddles(int L, DeviceBuffer& buf) {
    using T2 = typename vec2<T>::type;
    std::vector<T2> h(L);
    for (int m = 0; m < L; m++) {
        const double a = 2.0 * M_PI * (double)m / (double)L;
        h[m] = T2{(T)std::cos(a), (T)-std::sin(a)};
    }
    buf.allocate(sizeof(T2) * L);
    FTS_HIP_CHECK(hipMemcpy(buf.get(), h.data(), sizeof(T2) * L, hipMemcpyHostToDevice));
}

constexpr size_t LDS_LIMIT = 160 * 1024;

template <typename K>
void allow_lds(K kernel) {
    FTS_HIP_CHECK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_LIMIT));
}

template <typename T, int MAXR, bool BIG>
void allow_lds_slab() {
    allow_lds(&slab_kernel<T, SLAB_FWD, MAXR, BIG>);
    allow_lds(&slab_kernel<T, SLAB_INV1, MAXR, BIG>);
    allow_lds(&slab_kernel<T, SLAB_INV2, MAXR, BIG>);
}

template <typename T>
void set_lds_attributes() {
    static bool done = false;
    if (done) return;
    allow_lds_slab<T, 8, false>();
    allow_lds_slab<T, 16, false>();
    allow_lds_slab<T, 8, true>();
    allow_lds_slab<T, 16, true>();
    allow_lds(&xpass_kernel<T, false>);
    allow_lds(&xpass_kernel<T, true>);
    done = true;
}

template <int N>
void allow_lds_ct() {
    allow_lds(&slab_kernel_ct<double, SLAB_FWD, N>);
    allow_lds(&slab_kernel_ct<double, SLAB_INV1, N>);
    allow_lds(&slab_kernel_ct<double, SLAB_INV2, N>);
    allow_lds(&slab_kernel_ct<float, SLAB_FWD, N>);
    allow_lds(&slab_kernel_ct<float, SLAB_INV1, N>);
    allow_lds(&slab_kernel_ct<float, SLAB_INV2, N>);
    allow_lds(&xpass_kernel_ct<double, N>);
    allow_lds(&xpass_kernel_ct<float, N>);
}

// Launch-shape candidates of the wave-local kernels (the autotuner's search space)
constexpr bool slab_bd_ok(int N, int bd) {
    return (N == 32 && (bd == 128 || bd == 256)) || (N == 48 && (bd == 256 || bd == 384)) ||
           (N == 64 && (bd == 256 || bd == 320 || bd == 512)) || (N == 96 && (bd == 512 || bd == 1024)) ||
           (N == 128 && (bd == 512 || bd == 1024));
}
constexpr int SLAB_BD_ALL[] = {128, 256, 320, 384, 512, 1024};
constexpr int PENCIL_NL_ALL[] = {4, 8, 16};

template <typename T, int N, int BD>
void launch_slab_wl(const Grid& g, int mode, const SlabBatch<T>& b) {
    using G = WSlab<N, BD>;
    const dim3 grid(N, b.n), block(BD);
    const size_t lds = G::lds_elems * sizeof(typename vec2<T>::type);
    if (mode == SLAB_FWD)
        hipLaunchKernelGGL((slab_kernel_wl<T, SLAB_FWD, N, BD>), grid, block, lds, g.stream(), b, g.tw<T>(1),
                           g.tw<T>(2), g.sg.inv_m);
    else if (mode == SLAB_INV1)
        hipLaunchKernelGGL((slab_kernel_wl<T, SLAB_INV1, N, BD>), grid, block, lds, g.stream(), b, g.tw<T>(1),
                           g.tw<T>(2), g.sg.inv_m);
    else
        hipLaunchKernelGGL((slab_kernel_wl<T, SLAB_INV2, N, BD>), grid, block, lds, g.stream(), b, g.tw<T>(1),
                           g.tw<T>(2), g.sg.inv_m);
}

template <typename T, int N>
void launch_slab_wl_bd(const Grid& g, int bd, int mode, const SlabBatch<T>& b) {
    switch (bd) {
        case 128: if constexpr (slab_bd_ok(N, 128)) return launch_slab_wl<T, N, 128>(g, mode, b); break;
        case 256: if constexpr (slab_bd_ok(N, 256)) return launch_slab_wl<T, N, 256>(g, mode, b); break;
        case 320: if constexpr (slab_bd_ok(N, 320)) return launch_slab_wl<T, N, 320>(g, mode, b); break;
        case 384: if constexpr (slab_bd_ok(N, 384)) return launch_slab_wl<T, N, 384>(g, mode, b); break;
        case 512: if constexpr (slab_bd_ok(N, 512)) return launch_slab_wl<T, N, 512>(g, mode, b); break;
        case 1024: if constexpr (slab_bd_ok(N, 1024)) return launch_slab_wl<T, N, 1024>(g, mode, b); break;
        default: break;
    }
    throw Error("no wave-local slab kernel of " + std::to_string(bd) + " threads for N=" + std::to_string(N));
}

template <typename T, int N, int NL>
void launch_xpass_wl(const Grid& g, const XBatch<T>& b, const T* bx, const T* by, const T* bz) {
    using G = WPencil<N, NL>;
    const dim3 grid(N * wpencil_tiles<typename vec2<T>::type, N, NL>(), b.n), block(G::BD);
    hipLaunchKernelGGL((xpass_kernel_wl<T, N, NL>), grid, block, G::lds_elems * sizeof(typename vec2<T>::type),
                       g.stream(), b, g.tw<T>(0), bx, by, bz);
}

template <typename T, int N>
void launch_xpass_wl_nl(const Grid& g, int nl, const XBatch<T>& b, const T* bx, const T* by, const T* bz) {
    switch (nl) {
        case 4: return launch_xpass_wl<T, N, 4>(g, b, bx, by, bz);
        case 8: return launch_xpass_wl<T, N, 8>(g, b, bx, by, bz);
        case 16: return launch_xpass_wl<T, N, 16>(g, b, bx, by, bz);
        default: break;
    }
    throw Error("no wave-local pencil kernel of " + std::to_string(nl) + " columns");
}

template <typename T, int N, int NL>
int launch_stress_wl(const Grid& g, const StressBatch<T>& b, const StressGeom& sg, const T* bx, const T* by,
                     const T* bz, double* partials) {
    using G = WPencil<N, NL>;
    const dim3 grid(N * G::TILES, b.n), block(G::BD);
    hipLaunchKernelGGL((stress_pencil_wl<T, N, NL>), grid, block, G::lds_elems * sizeof(typename vec2<T>::type),
                       g.stream(), b, sg, g.tw<T>(0), bx, by, bz, partials);
    return N * G::TILES;
}

template <typename T, int N>
int launch_stress_wl_nl(const Grid& g, int nl, const StressBatch<T>& b, const StressGeom& sg, const T* bx,
                        const T* by, const T* bz, double* partials) {
    switch (nl) {
        case 4: return launch_stress_wl<T, N, 4>(g, b, sg, bx, by, bz, partials);
        case 8: return launch_stress_wl<T, N, 8>(g, b, sg, bx, by, bz, partials);
        case 16: return launch_stress_wl<T, N, 16>(g, b, sg, bx, by, bz, partials);
        default: break;
    }
    throw Error("no wave-local stress kernel of " + std::to_string(nl) + " columns");
}

template <typename T>
constexpr int prec_index() {
    return std::is_same<T, float>::value ? 1 : 0;
}

// LDS limits of every wave-local variant, set before any launch (and so before any
// graph capture could record the first one)
template <typename T, int N, int BD>
void allow_lds_slab_wl() {
    if constexpr (slab_bd_ok(N, BD)) {
        allow_lds(&slab_kernel_wl<T, SLAB_FWD, N, BD>);
        allow_lds(&slab_kernel_wl<T, SLAB_INV1, N, BD>);
        allow_lds(&slab_kernel_wl<T, SLAB_INV2, N, BD>);
    }
}

template <typename T, int N>
void allow_lds_wl() {
    allow_lds_slab_wl<T, N, 128>();
    allow_lds_slab_wl<T, N, 256>();
    allow_lds_slab_wl<T, N, 320>();
    allow_lds_slab_wl<T, N, 384>();
    allow_lds_slab_wl<T, N, 512>();
    allow_lds_slab_wl<T, N, 1024>();
    allow_lds(&xpass_kernel_wl<T, N, 4>);
    allow_lds(&xpass_kernel_wl<T, N, 8>);
    allow_lds(&xpass_kernel_wl<T, N, 16>);
    allow_lds(&stress_pencil_wl<T, N, 4>);
    allow_lds(&stress_pencil_wl<T, N, 8>);
    allow_lds(&stress_pencil_wl<T, N, 16>);
}

template <int N>
void allow_lds_n() {
    allow_lds_ct<N>();
    allow_lds_wl<double, N>();
    allow_lds_wl<float, N>();
}

void set_lds_attributes_ct() {
    static bool done = false;
    if (done) return;
    allow_lds_n<32>();
    allow_lds_n<48>();
    allow_lds_n<64>();
    allow_lds_n<96>();
    allow_lds_n<128>();
    done = true;
}

template <typename T, int N>
void launch_slab_ct(const Grid& g, int mode, const SlabBatch<T>& b) {
    if (g.wave_local) {
        if constexpr (N > 64) {
            // at these sizes the wave-local kernel parks the Richardson q1 in the output
            // slice and starts a second forward transform from qin / qout
            for (int i = 0; i < b.n; i++) {
                const SlabItem<T>& it = b.it[i];
                if ((mode == SLAB_INV2 && !it.qout) || (mode == SLAB_FWD && it.n_out > 1 && !it.qin))
                    throw Error("slab item without the real slice its second pass reads (N=" + std::to_string(N) + ")");
            }
        }
        const int bd_mode = g.slab_bd_mode[prec_index<T>()][mode];
        const int bd = bd_mode ? bd_mode : g.slab_bd[prec_index<T>()];
        launch_slab_wl_bd<T, N>(g, bd ? bd : wslab_default_bd(N), mode, b);
        return;
    }
    using G = SlabShape<N>;
    const dim3 grid(N, b.n), block(G::BD);
    const size_t lds = G::lds_elems * sizeof(typename vec2<T>::type);
    if (mode == SLAB_FWD)
        hipLaunchKernelGGL((slab_kernel_ct<T, SLAB_FWD, N>), grid, block, lds, g.stream(), b, g.tw<T>(1), g.tw<T>(2),
                           g.sg.inv_m);
    else if (mode == SLAB_INV1)
        hipLaunchKernelGGL((slab_kernel_ct<T, SLAB_INV1, N>), grid, block, lds, g.stream(), b, g.tw<T>(1), g.tw<T>(2),
                           g.sg.inv_m);
    else
        hipLaunchKernelGGL((slab_kernel_ct<T, SLAB_INV2, N>), grid, block, lds, g.stream(), b, g.tw<T>(1), g.tw<T>(2),
                           g.sg.inv_m);
}

template <typename T, int N>
void launch_xpass_ct(const Grid& g, const XBatch<T>& b, const T* bx, const T* by, const T* bz) {
    if (g.wave_local) {
        const int nl = g.pencil_nl[prec_index<T>()];
        launch_xpass_wl_nl<T, N>(g, nl ? nl : wpencil_default_nl(N, (int)sizeof(typename vec2<T>::type)), b, bx,
                                 by, bz);
        return;
    }
    using G = PencilShape<N>;
    const dim3 grid(N * G::TILES, b.n), block(G::BD);
    hipLaunchKernelGGL((xpass_kernel_ct<T, N>), grid, block, G::lds_elems * sizeof(typename vec2<T>::type), g.stream(),
                       b, g.tw<T>(0), bx, by, bz);
}

template <typename T, int MODE, int MAXR, bool BIG>
void launch_slab_k(const Grid& g, const SlabBatch<T>& b, dim3 grid, dim3 block, size_t lds) {
    hipLaunchKernelGGL((slab_kernel<T, MODE, MAXR, BIG>), grid, block, lds, g.stream(), g.sg, b, g.tw<T>(1),
                       g.tw<T>(2));
}

template <typename T, int MODE>
void launch_slab_mode(const Grid& g, const SlabBatch<T>& b, dim3 grid, dim3 block, size_t lds) {
    if (g.slab_maxr == 8) {
        if (g.big) launch_slab_k<T, MODE, 8, true>(g, b, grid, block, lds);
        else launch_slab_k<T, MODE, 8, false>(g, b, grid, block, lds);
    } else {
        if (g.big) launch_slab_k<T, MODE, 16, true>(g, b, grid, block, lds);
        else launch_slab_k<T, MODE, 16, false>(g, b, grid, block, lds);
    }
}

}  // namespace

Grid::Grid(fts_context* c, int d, const int* n, const double* l, const double* mask) : ctx(c), dim(d) {
    FTS_REQUIRE(dim >= 1 && dim <= 3, "dim must be 1, 2 or 3");
    M = 1;
    for (int i = 0; i < dim; i++) {
        FTS_REQUIRE(n[i] > 0, "nx must be positive");
        FTS_REQUIRE(l[i] > 0, "lx must be positive");
        nx[i] = n[i];
        lx[i] = l[i];
        dx[i] = l[i] / n[i];
        M *= n[i];
    }
    for (int i = dim; i < 3; i++) {
        nx[i] = 1;
        lx[i] = 1.0;
        dx[i] = 1.0;
    }
    for (int a = 0; a < 3; a++) {
        const int src = a - (3 - dim);
        tnx[a] = src >= 0 ? nx[src] : 1;
        tdx[a] = src >= 0 ? dx[src] : 1.0;
    }
    nzc = tnx[2] / 2 + 1;
    Mc = tnx[0] * tnx[1] * nzc;
    spitch = spitch32 = nzc;
    Ms = (size_t)Mc;
    // ComputationBox.cpp:47-98: dv = prod dx (times mask), volume = prod lx
    dV = 1.0;
    for (int i = 0; i < dim; i++) dV *= dx[i];
    volume = 1.0;
    for (int i = 0; i < dim; i++) volume *= lx[i];
    if (mask) {
        mask_h.resize(M);
        std::vector<float> m32(M);
        for (int i = 0; i < M; i++) {
            if (std::abs(mask[i]) < 1e-7) mask_h[i] = 0.0;
            else if (std::abs(mask[i] - 1.0) < 1e-7) mask_h[i] = 1.0;
            else throw Error("mask[" + std::to_string(i) + "] must be 0.0 or 1.0");
            m32[i] = (float)mask_h[i];
        }
        mask64.allocate(sizeof(double) * M);
        mask32.allocate(sizeof(float) * M);
        FTS_HIP_CHECK(hipMemcpy(mask64.get(), mask_h.data(), sizeof(double) * M, hipMemcpyHostToDevice));
        FTS_HIP_CHECK(hipMemcpy(mask32.get(), m32.data(), sizeof(float) * M, hipMemcpyHostToDevice));
    }

    partials.allocate(sizeof(double) * MAX_BATCH * RED_BLOCKS * 3);
    scalars.allocate(sizeof(double) * MAX_BATCH);
    scratch64.allocate(sizeof(double) * 2 * (size_t)M);
    // the FFT plans serve the pseudo-spectral solvers only; a grid they cannot
    // transform still serves the real-space solver, so the failure is kept for
    // the first FFT use
    try {
        setup_fft();
    } catch (const Error& e) {
        fft_error = e.what();
    }
}

void Grid::setup_fft() {
    ax = plan_axis(tnx[0]);
    ay = plan_axis(tnx[1]);
    az = plan_axis(tnx[2]);
    for (int a = 0; a < 3; a++) {
        upload_twiddles<double>(tnx[a], tw64[a]);
        upload_twiddles<float>(tnx[a], tw32[a]);
    }

    sg.ny = tnx[1];
    sg.nz = tnx[2];
    sg.nzc = nzc;
    sg.P = nzc;
    sg.rows = 2 * ((tnx[1] + 1) / 2);
    sg.ay = ay;
    sg.az = az;
    sg.inv_m = 1.0 / (double)M;
    slab_threads = choose_threads(sg.ny * sg.nz);
    const int per_thread = ceil_div(sg.ny * sg.nz, slab_threads);
    FTS_REQUIRE(per_thread <= 16, "y-z slab too large (more than 16384 points)");
    // the 8-value variant is compiled for at most 512 threads (more registers per thread)
    slab_maxr = (per_thread <= 8 && slab_threads <= 512) ? 8 : 16;
    sg.ay.group = group_lines(sg.ay, slab_threads);
    sg.az.group = group_lines(sg.az, slab_threads);
    FTS_REQUIRE(slab_lds<double>() <= LDS_LIMIT, "y-z slab does not fit in LDS");
    big = plan_is_big(ax) || plan_is_big(ay) || plan_is_big(az);

    xg.nx = tnx[0];
    xg.ny = tnx[1];
    xg.nzc = nzc;
    int C = nzc, tiles = 1;
    while (tnx[1] * tiles < 128 && C > 8) {
        C = ceil_div(C, 2);
        tiles = ceil_div(nzc, C);
    }
    while ((size_t)(tnx[0] * C + tnx[0]) * sizeof(double2) > 64 * 1024 && C > 1) {
        C = ceil_div(C, 2);
        tiles = ceil_div(nzc, C);
    }
    xg.C = C;
    xg.n_tiles = tiles;
    xg.P = C;
    xg.ax = ax;
    x_threads = choose_threads(tnx[0] * C);
    xg.ax.group = group_lines(xg.ax, x_threads);

    set_lds_attributes<double>();
    set_lds_attributes<float>();
    // cubic grids with a compile-time specialised kernel pair (env FTS_NO_CT=1 keeps the runtime kernels)
    ct_n = 0;
    if (dim == 3 && tnx[0] == tnx[1] && tnx[1] == tnx[2] && !std::getenv("FTS_NO_CT")) {
        const int n = tnx[0];
        if (n == 32 || n == 48 || n == 64 || n == 96 || n == 128) ct_n = n;
    }
    // wave-local line transforms for those grids (env FTS_CT_KERNELS=1: the block-synchronous ones)
    wave_local = ct_n != 0 && !std::getenv("FTS_CT_KERNELS");
    if (ct_n) set_lds_attributes_ct();
    if (wave_local) {
        spitch = wspec_pitch<vec2<double>::type>(ct_n);
        spitch32 = wspec_pitch<vec2<float>::type>(ct_n);
        Ms = (size_t)tnx[0] * tnx[1] * std::max(spitch, spitch32);
        tune_fft_schedule<double>();
        tune_fft_schedule<float>();
    }
}

// ------------------------------------------------------------------ FFT schedule autotuner
// Times every pair of launch shapes of the wave-local slab and pencil kernels for this
// grid and precision on scratch fields (steps of two propagators: slab inverse+forward
// with the exp(-w ds) factors, then the pencil pass with the bond factors, alternating
// as in a propagator), and keeps the fastest pair; it replaces the defaults only when
// at least 2% faster.
// The choice is cached per (N, precision) for the process. FTS_FFT_TUNE=0 keeps the
// defaults. Candidates run the same per-line transforms; results agree to rounding
// (bit-identical at 32^3 and 64^3, where the compiler contracts both alike).
namespace {
struct TuneChoice {
    int bd, nl;
    std::string note;
    int bd_mode[3];   // per slab mode, 0: bd
};
std::map<std::pair<int, int>, TuneChoice>& tune_cache() {
    static std::map<std::pair<int, int>, TuneChoice> c;
    return c;
}
std::mutex& tune_mutex() {   // grids created on several threads tune one (N, precision) once
    static std::mutex m;
    return m;
}

std::string us_text(int bd, int nl, float ms) {
    char buf[48];
    std::snprintf(buf, sizeof(buf), " %d/%d:%.2f", bd, nl, 1e3 * ms);
    return buf;
}

template <typename F>
float time_launches(hipStream_t st, F&& launch) {
    hipEvent_t e0, e1;
    FTS_HIP_CHECK(hipEventCreate(&e0));
    FTS_HIP_CHECK(hipEventCreate(&e1));
    for (int i = 0; i < 3; i++) launch();
    float best = 1e30f;
    for (int rep = 0; rep < 3; rep++) {
        FTS_HIP_CHECK(hipEventRecord(e0, st));
        for (int i = 0; i < 8; i++) launch();
        FTS_HIP_CHECK(hipEventRecord(e1, st));
        FTS_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        FTS_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, ms / 8.f);
    }
    FTS_HIP_CHECK(hipGetLastError());
    FTS_HIP_CHECK(hipEventDestroy(e0));
    FTS_HIP_CHECK(hipEventDestroy(e1));
    return best;
}

template <typename T, int N>
TuneChoice tune_wave_local(const Grid& g) {
    using T2 = typename vec2<T>::type;
    constexpr int NZC = N / 2 + 1, ITEMS = 2;
    const size_t M = (size_t)N * N * N, Mc = (size_t)N * N * wspec_pitch<T2>(N);
    DeviceBuffer spec, q, w, tab;
    // two spectra per item (c0, c1: the Richardson step and the two-spectrum forward
    // transform of continuous chains) and one real slice (qin and qout)
    spec.allocate(sizeof(T2) * Mc * 2 * ITEMS);
    q.allocate(sizeof(T) * M * ITEMS);
    w.allocate(sizeof(T) * M);
    tab.allocate(sizeof(T) * (2 * N + NZC));
    hipStream_t st = g.stream();
    FTS_HIP_CHECK(hipMemsetAsync(spec.get(), 0, sizeof(T2) * Mc * 2 * ITEMS, st));
    FTS_HIP_CHECK(hipMemsetAsync(q.get(), 0, sizeof(T) * M * ITEMS, st));
    FTS_HIP_CHECK(hipMemsetAsync(w.get(), 0, sizeof(T) * M, st));
    FTS_HIP_CHECK(hipMemsetAsync(tab.get(), 0, sizeof(T) * (2 * N + NZC), st));
    SlabBatch<T> sb{};
    XBatch<T> xb{};
    sb.n = xb.n = ITEMS;
    for (int i = 0; i < ITEMS; i++) {
        SlabItem<T>& it = sb.it[i];
        it.qin = q.as<T>() + i * M;
        it.qout = q.as<T>() + i * M;
        it.c0 = spec.as<T2>() + 2 * i * Mc;
        it.c1 = spec.as<T2>() + (2 * i + 1) * Mc;
        it.w = w.as<T>();
        // FP64 discrete steps read exp(-w ds) from a table (coef < 0), FP32 steps evaluate it
        it.post0 = it.pre0 = std::is_same<T, double>::value ? (T)-1 : (T)0.01;
        it.post1 = it.pre1 = (T)0.01;
        it.n_out = 1;
        it.forward = 1;
        xb.it[i].c = spec.as<T2>() + 2 * i * Mc;
        xb.it[i].table = 0;
    }
    const T* bx = tab.as<T>();
    TuneChoice c{wslab_default_bd(N), wpencil_default_nl(N, (int)sizeof(T2)), "", {0, 0, 0}};
    // whole steps (slab pass then pencil pass, as the propagators alternate them)
    std::string note = "us per step of 2 propagators, slab threads/pencil columns";
    float t_def = 0.f, t_best = 1e30f;
    TuneChoice best = c;
    for (int bd : SLAB_BD_ALL) {
        if (!slab_bd_ok(N, bd)) continue;
        for (int nl : PENCIL_NL_ALL) {
            const float t = time_launches(st, [&] {
                launch_slab_wl_bd<T, N>(g, bd, SLAB_INV1, sb);
                launch_xpass_wl_nl<T, N>(g, nl, xb, bx, bx + N, bx + 2 * N);
            });
            note += us_text(bd, nl, t);
            if (bd == c.bd && nl == c.nl) t_def = t;
            if (t < t_best) {
                t_best = t;
                best.bd = bd;
                best.nl = nl;
            }
        }
    }
    if (t_best < 0.98f * t_def) {
        c.bd = best.bd;
        c.nl = best.nl;
    }
    // the forward (two spectra out) and Richardson modes of continuous chains hold
    // more values per thread than the half step: each picks its own workgroup size
    // among the candidates, timed alone (the pencil pass does not depend on it)
    SlabBatch<T> sc = sb;
    for (int i = 0; i < ITEMS; i++) {
        sc.it[i].post0 = sc.it[i].pre0 = (T)0.01;
        sc.it[i].n_out = 2;
    }
    for (int mode : {SLAB_FWD, SLAB_INV2}) {
        note += mode == SLAB_FWD ? "; forward, two spectra:" : "; Richardson:";
        float t_c = 0.f, t_m = 1e30f;
        int bd_m = c.bd;
        for (int bd : SLAB_BD_ALL) {
            if (!slab_bd_ok(N, bd)) continue;
            const float t = time_launches(st, [&] { launch_slab_wl_bd<T, N>(g, bd, mode, sc); });
            char buf[32];
            std::snprintf(buf, sizeof(buf), " %d:%.2f", bd, 1e3 * t);
            note += buf;
            if (bd == c.bd) t_c = t;
            if (t < t_m) {
                t_m = t;
                bd_m = bd;
            }
        }
        if (bd_m != c.bd && t_m < 0.98f * t_c) c.bd_mode[mode] = bd_m;
    }
    FTS_HIP_CHECK(hipStreamSynchronize(st));
    c.note = "tuned (" + note + ")";
    return c;
}
}  // namespace

template <typename T>
void Grid::tune_fft_schedule() {
    const int p = prec_index<T>();
    slab_bd[p] = wslab_default_bd(ct_n);
    for (int m = 0; m < 3; m++) slab_bd_mode[p][m] = 0;
    pencil_nl[p] = wpencil_default_nl(ct_n, (int)sizeof(typename vec2<T>::type));
    // Tuning picks launch shapes by wall-clock timing, so the choice can differ between
    // processes. At 32^3 and 64^3 every candidate gives bit-identical results; at 48^3,
    // 96^3 and 128^3 they differ by rounding, and a run could then follow another
    // trajectory in another process: there the defaults stay unless FTS_FFT_TUNE=1.
    const char* env = std::getenv("FTS_FFT_TUNE");
    const bool identical = ct_n == 32 || ct_n == 64;
    if (!wave_local || (env && env[0] == '0')) {
        fft_schedule[p] = "defaults";
        return;
    }
    if (!identical && !(env && env[0] == '1')) {
        fft_schedule[p] = "defaults (candidates at this size differ by rounding; FTS_FFT_TUNE=1 tunes)";
        return;
    }
    std::lock_guard<std::mutex> lock(tune_mutex());
    auto& cache = tune_cache();
    const auto key = std::make_pair(ct_n, p);
    auto hit = cache.find(key);
    if (hit == cache.end()) {
        TuneChoice c{slab_bd[p], pencil_nl[p], "", {0, 0, 0}};
        try {
            switch (ct_n) {
                case 32: c = tune_wave_local<T, 32>(*this); break;
                case 48: c = tune_wave_local<T, 48>(*this); break;
                case 64: c = tune_wave_local<T, 64>(*this); break;
                case 96: c = tune_wave_local<T, 96>(*this); break;
                default: c = tune_wave_local<T, 128>(*this); break;
            }
        } catch (const Error& e) {
            (void)hipGetLastError();
            c = TuneChoice{wslab_default_bd(ct_n), wpencil_default_nl(ct_n, (int)sizeof(typename vec2<T>::type)),
                           std::string("defaults (tuning failed: ") + e.what() + ")", {0, 0, 0}};
        }
        hit = cache.emplace(key, c).first;
    }
    slab_bd[p] = hit->second.bd;
    for (int m = 0; m < 3; m++) slab_bd_mode[p][m] = hit->second.bd_mode[m];
    pencil_nl[p] = hit->second.nl;
    fft_schedule[p] = hit->second.note;
}

void Grid::set_fft_schedule(int precision, int threads, int columns) {
    FTS_REQUIRE(wave_local, "this grid has no wave-local FFT kernels (cubic 32, 48, 64, 96 or 128 only)");
    FTS_REQUIRE(precision == 0 || precision == 1, "precision must be FTS_FP64 or FTS_FP32");
    FTS_REQUIRE(slab_bd_ok(ct_n, threads), "no slab kernel of " + std::to_string(threads) + " threads for N=" +
                                               std::to_string(ct_n));
    FTS_REQUIRE(columns == 4 || columns == 8 || columns == 16, "pencil columns must be 4, 8 or 16");
    slab_bd[precision] = threads;
    for (int m = 0; m < 3; m++) slab_bd_mode[precision][m] = 0;
    pencil_nl[precision] = columns;
    fft_schedule[precision] = "set by the caller";
}

// ComputationBox.cpp:229-262
void Grid::set_lx(const double* l) {
    for (int i = 0; i < dim; i++) FTS_REQUIRE(l[i] > 0, "new lx must be positive");
    for (int i = 0; i < dim; i++) {
        lx[i] = l[i];
        dx[i] = l[i] / nx[i];
    }
    for (int a = 0; a < 3; a++) {
        const int src = a - (3 - dim);
        tdx[a] = src >= 0 ? dx[src] : 1.0;
    }
    dV = 1.0;
    for (int i = 0; i < dim; i++) dV *= dx[i];
    volume = 0.0;
    for (int i = 0; i < M; i++) volume += mask_h.empty() ? dV : dV * mask_h[i];
}

template <>
const double* Grid::mask<double>() const {
    return mask64.as<double>();
}
template <>
const float* Grid::mask<float>() const {
    return mask32.as<float>();
}
template <>
const double2* Grid::tw<double>(int a) const {
    return tw64[a].as<double2>();
}
template <>
const float2* Grid::tw<float>(int a) const {
    return tw32[a].as<float2>();
}

// The FP32 variants with the generic odd radices (7..31) hung on the GPU with no
// identified cause; they are switched off until one is found.
template <typename T>
static void require_supported(const Grid& g) {
    if (std::is_same<T, float>::value && g.big)
        throw Error("FP32 propagators need grid sizes whose prime factors are 2, 3 and 5");
}

// algorithmic bytes of one slab launch: spectra, fields and w each moved once
template <typename T>
static double slab_bytes(const Grid& g, int mode, const SlabBatch<T>& b) {
    const double e = sizeof(T), spec = 2.0 * e * g.Mc, field = e * g.M;
    double bytes = 0.0;
    for (int i = 0; i < b.n; i++) {
        const SlabItem<T>& it = b.it[i];
        if (it.w) bytes += field;
        if (it.mask) bytes += field;
        if (mode == SLAB_FWD) {
            bytes += field + it.n_out * spec;
        } else {
            const int n_spec = mode == SLAB_INV2 ? 2 : 1;
            bytes += n_spec * spec;
            if (it.qout) bytes += field;
            if (it.forward) bytes += n_spec * spec;
        }
    }
    return bytes;
}

template <typename T>
void Grid::launch_slab(int mode, const SlabBatch<T>& b) const {
    require_supported<T>(*this);
    if (b.n == 0) return;
    KernelTimer& timer = ctx->timer;
    hipEvent_t start = timer.begin(stream(), FTS_KERNEL_SLAB);
    switch (ct_n) {
        case 32: launch_slab_ct<T, 32>(*this, mode, b); break;
        case 48: launch_slab_ct<T, 48>(*this, mode, b); break;
        case 64: launch_slab_ct<T, 64>(*this, mode, b); break;
        case 96: launch_slab_ct<T, 96>(*this, mode, b); break;
        case 128: launch_slab_ct<T, 128>(*this, mode, b); break;
        default: {
            const dim3 grid(tnx[0], b.n), block(slab_threads);
            const size_t lds = slab_lds<T>();
            if (mode == SLAB_FWD) launch_slab_mode<T, SLAB_FWD>(*this, b, grid, block, lds);
            else if (mode == SLAB_INV1) launch_slab_mode<T, SLAB_INV1>(*this, b, grid, block, lds);
            else launch_slab_mode<T, SLAB_INV2>(*this, b, grid, block, lds);
        }
    }
    FTS_HIP_CHECK(hipGetLastError());
    if (start) timer.end(stream(), start, FTS_KERNEL_SLAB, slab_bytes(*this, mode, b));
}

template <typename T>
void Grid::launch_xpass(const XBatch<T>& b, const T* bx, const T* by, const T* bz) const {
    require_supported<T>(*this);
    if (b.n == 0) return;
    KernelTimer& timer = ctx->timer;
    hipEvent_t start = timer.begin(stream(), FTS_KERNEL_PENCIL);
    switch (ct_n) {
        case 32: launch_xpass_ct<T, 32>(*this, b, bx, by, bz); break;
        case 48: launch_xpass_ct<T, 48>(*this, b, bx, by, bz); break;
        case 64: launch_xpass_ct<T, 64>(*this, b, bx, by, bz); break;
        case 96: launch_xpass_ct<T, 96>(*this, b, bx, by, bz); break;
        case 128: launch_xpass_ct<T, 128>(*this, b, bx, by, bz); break;
        default: {
            const dim3 grid(tnx[1] * xg.n_tiles, b.n), block(x_threads);
            if (big)
                hipLaunchKernelGGL((xpass_kernel<T, true>), grid, block, x_lds<T>(), stream(), xg, b, tw<T>(0), bx,
                                   by, bz);
            else
                hipLaunchKernelGGL((xpass_kernel<T, false>), grid, block, x_lds<T>(), stream(), xg, b, tw<T>(0), bx,
                                   by, bz);
        }
    }
    FTS_HIP_CHECK(hipGetLastError());
    // each spectrum read and written once
    if (start) timer.end(stream(), start, FTS_KERNEL_PENCIL, 4.0 * sizeof(T) * Mc * b.n);
}

template <typename T>
int Grid::launch_stress_pencil(const StressBatch<T>& b, const StressGeom& sg, const T* bx, const T* by, const T* bz,
                               double* partials) const {
    require_supported<T>(*this);
    if (!wave_local) throw Error("the fused stress pencil pass needs a wave-local grid");
    // tiles of one 128-byte line per spectrum row (rows start on a line): narrower tiles
    // share lines with their neighbours, which run on other XCDs with their own L2, and
    // every line is then fetched twice
    const int nl = 128 / (int)sizeof(typename vec2<T>::type);
    int blocks = 0;
    switch (ct_n) {
        case 32: blocks = launch_stress_wl_nl<T, 32>(*this, nl, b, sg, bx, by, bz, partials); break;
        case 48: blocks = launch_stress_wl_nl<T, 48>(*this, nl, b, sg, bx, by, bz, partials); break;
        case 64: blocks = launch_stress_wl_nl<T, 64>(*this, nl, b, sg, bx, by, bz, partials); break;
        case 96: blocks = launch_stress_wl_nl<T, 96>(*this, nl, b, sg, bx, by, bz, partials); break;
        case 128: blocks = launch_stress_wl_nl<T, 128>(*this, nl, b, sg, bx, by, bz, partials); break;
        default: throw Error("the fused stress pencil pass needs an N^3 grid of a specialised size");
    }
    FTS_HIP_CHECK(hipGetLastError());
    return blocks;
}

template <typename T>
void Grid::launch_dots(const DotBatch<T>& b, double* out) const {
    if (b.n == 0) return;
    hipLaunchKernelGGL((dot_stage1<T>), dim3(RED_BLOCKS, b.n), dim3(RED_THREADS), 0, stream(), b, M, dV,
                       mask_h.empty() ? nullptr : mask64.as<double>(), partials.as<double>());
    FTS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL((dot_stage2<T>), dim3(b.n), dim3(RED_BLOCKS), 0, stream(), b, partials.as<double>(), out);
    FTS_HIP_CHECK(hipGetLastError());
}

double Grid::inner_product(const double* a, const double* b, bool on_device) {
    const double* da = a;
    const double* db = b;
    if (!on_device) {
        FTS_HIP_CHECK(hipMemcpyAsync(scratch64.as<double>(), a, sizeof(double) * M, hipMemcpyHostToDevice, stream()));
        FTS_HIP_CHECK(hipMemcpyAsync(scratch64.as<double>() + M, b, sizeof(double) * M, hipMemcpyHostToDevice,
                                     stream()));
        da = scratch64.as<double>();
        db = scratch64.as<double>() + M;
    }
    DotBatch<double> batch{};
    batch.n = 1;
    batch.it[0] = DotItem<double>{da, db, nullptr, 0.0, 1.0};
    launch_dots(batch, scalars.as<double>());
    double r;
    FTS_HIP_CHECK(hipMemcpyAsync(&r, scalars.as<double>(), sizeof(double), hipMemcpyDeviceToHost, stream()));
    FTS_HIP_CHECK(hipStreamSynchronize(stream()));
    return r;
}

double Grid::integral(const double* a, bool on_device) {
    const double* da = a;
    if (!on_device) {
        FTS_HIP_CHECK(hipMemcpyAsync(scratch64.as<double>(), a, sizeof(double) * M, hipMemcpyHostToDevice, stream()));
        da = scratch64.as<double>();
    }
    DotBatch<double> batch{};
    batch.n = 1;
    batch.it[0] = DotItem<double>{da, nullptr, nullptr, 0.0, 1.0};
    launch_dots(batch, scalars.as<double>());
    double r;
    FTS_HIP_CHECK(hipMemcpyAsync(&r, scalars.as<double>(), sizeof(double), hipMemcpyDeviceToHost, stream()));
    FTS_HIP_CHECK(hipStreamSynchronize(stream()));
    return r;
}

template <typename T>
void Grid::fft_forward(const T* in, typename vec2<T>::type* out, int batch) {
    require_fft();
    using T2 = typename vec2<T>::type;
    // padded rows: transform into a workspace of the padded layout, then copy the rows out
    const int pitch = spec_pitch<T>();
    const bool padded = pitch != nzc;
    DeviceBuffer tmp;
    if (padded) tmp.allocate(sizeof(T2) * Ms * std::min(batch, MAX_BATCH));
    for (int b0 = 0; b0 < batch; b0 += MAX_BATCH) {
        const int nb = std::min(MAX_BATCH, batch - b0);
        SlabBatch<T> sb{};
        XBatch<T> xb{};
        sb.n = xb.n = nb;
        for (int b = 0; b < nb; b++) {
            T2* spec = padded ? tmp.as<T2>() + (size_t)b * Ms : out + (size_t)(b0 + b) * Mc;
            SlabItem<T>& it = sb.it[b];
            it.qin = in + (size_t)(b0 + b) * M;
            it.c0 = spec;
            it.n_out = 1;
            xb.it[b] = XItem<T>{spec, -1};
        }
        launch_slab<T>(SLAB_FWD, sb);
        launch_xpass<T>(xb, nullptr, nullptr, nullptr);
        for (int b = 0; padded && b < nb; b++)
            FTS_HIP_CHECK(hipMemcpy2DAsync(out + (size_t)(b0 + b) * Mc, sizeof(T2) * nzc, tmp.as<T2>() + (size_t)b * Ms,
                                           sizeof(T2) * pitch, sizeof(T2) * nzc, (size_t)tnx[0] * tnx[1],
                                           hipMemcpyDeviceToDevice, stream()));
    }
    if (padded) FTS_HIP_CHECK(hipStreamSynchronize(stream()));
}

template <typename T>
void Grid::fft_backward_ws(const typename vec2<T>::type* in, T* out, int batch, typename vec2<T>::type* ws,
                           int ws_fields) {
    require_fft();
    // the pencil pass transforms in place, so the input is first copied into the
    // workspace (ws_fields spectra of Ms elements), in chunks of at most ws_fields
    using T2 = typename vec2<T>::type;
    FTS_REQUIRE(ws_fields >= 1, "the FFT workspace holds no spectrum");
    const int chunk = std::min(ws_fields, MAX_BATCH);
    for (int b0 = 0; b0 < batch; b0 += chunk) {
        const int nb = std::min(chunk, batch - b0);
        // the caller's rows of nzc values into the workspace's rows of this precision's
        // pitch, one spectrum per slot of Ms elements (Ms fits the wider pitch of the two)
        for (int b = 0; b < nb; b++)
            FTS_HIP_CHECK(hipMemcpy2DAsync(ws + (size_t)b * Ms, sizeof(T2) * spec_pitch<T>(),
                                           in + (size_t)(b0 + b) * Mc, sizeof(T2) * nzc, sizeof(T2) * nzc,
                                           (size_t)tnx[0] * tnx[1], hipMemcpyDeviceToDevice, stream()));
        SlabBatch<T> sb{};
        XBatch<T> xb{};
        sb.n = xb.n = nb;
        for (int b = 0; b < nb; b++) {
            xb.it[b] = XItem<T>{ws + (size_t)b * Ms, -2};
            SlabItem<T>& it = sb.it[b];
            it.c0 = ws + (size_t)b * Ms;
            it.qout = out + (size_t)(b0 + b) * M;
        }
        launch_xpass<T>(xb, nullptr, nullptr, nullptr);
        launch_slab<T>(SLAB_INV1, sb);
    }
}

template <typename T>
void Grid::fft_backward(const typename vec2<T>::type* in, T* out, int batch) {
    require_fft();
    using T2 = typename vec2<T>::type;
    const int nw = std::min(batch, MAX_BATCH);
    DeviceBuffer tmp(sizeof(T2) * Ms * nw);
    fft_backward_ws<T>(in, out, batch, tmp.as<T2>(), nw);
    FTS_HIP_CHECK(hipStreamSynchronize(stream()));
}

// ------------------------------------------------------------------ no-allocation workspace
// fts_grid_fft_backward_noalloc's workspace: a fixed array in device memory, one copy
// per GPU, held until the process ends. Host code never passes `scratch` itself (its
// name is a host-side stand-in there); it uses the device address that
// hipGetSymbolAddress returns for the current device, kept per device below.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_ptr[MAX_DEVICES] = {};
}  // namespace

void noalloc_scratch_init(int device) {
    FTS_REQUIRE(device >= 0 && device < MAX_DEVICES, "device index beyond the no-allocation workspace table");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (scratch_ptr[device]) return;
    int cur = -1;
    FTS_HIP_CHECK(hipGetDevice(&cur));
    FTS_REQUIRE(cur == device, "no-allocation workspace set-up on a device that is not current");
    void* p = nullptr;
    FTS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    FTS_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address for the no-allocation workspace");
    scratch_ptr[device] = p;
}

template <typename T>
void Grid::fft_backward_noalloc(const typename vec2<T>::type* in, T* out, int batch) {
    require_fft();
    using T2 = typename vec2<T>::type;
    void* ws = nullptr;
    {
        std::lock_guard<std::mutex> lock(scratch_mu);
        ws = ctx->device >= 0 && ctx->device < MAX_DEVICES ? scratch_ptr[ctx->device] : nullptr;
    }
    FTS_REQUIRE(ws != nullptr, "the no-allocation workspace of this device is not set up");
    const size_t per_field = sizeof(T2) * Ms;
    if (per_field > NOALLOC_SCRATCH_BYTES)
        throw Error("fts_grid_fft_backward_noalloc: one spectrum of this grid takes " + std::to_string(per_field) +
                    " bytes of workspace; the no-allocation workspace holds " +
                    std::to_string(NOALLOC_SCRATCH_BYTES) + " bytes");
    fft_backward_ws<T>(in, out, batch, static_cast<T2*>(ws), (int)(NOALLOC_SCRATCH_BYTES / per_field));
}

template void Grid::launch_slab<double>(int, const SlabBatch<double>&) const;
template void Grid::launch_slab<float>(int, const SlabBatch<float>&) const;
template void Grid::launch_xpass<double>(const XBatch<double>&, const double*, const double*, const double*) const;
template void Grid::launch_xpass<float>(const XBatch<float>&, const float*, const float*, const float*) const;
template int Grid::launch_stress_pencil<double>(const StressBatch<double>&, const StressGeom&, const double*,
                                                const double*, const double*, double*) const;
template int Grid::launch_stress_pencil<float>(const StressBatch<float>&, const StressGeom&, const float*,
                                               const float*, const float*, double*) const;
template void Grid::launch_dots<double>(const DotBatch<double>&, double*) const;
template void Grid::launch_dots<float>(const DotBatch<float>&, double*) const;
template void Grid::fft_forward<double>(const double*, double2*, int);
template void Grid::fft_forward<float>(const float*, float2*, int);
template void Grid::fft_backward<double>(const double2*, double*, int);
template void Grid::fft_backward<float>(const float2*, float*, int);
template void Grid::fft_backward_noalloc<double>(const double2*, double*, int);
template void Grid::fft_backward_noalloc<float>(const float2*, float*, int);

// ------------------------------------------------------------------ bond tables
void BondTables::build(const Grid& g, const std::vector<double>& bsq, double ds) {
    const double PI = 3.14159265358979323846;
    const int nt = 2 * (int)bsq.size();
    std::vector<double> hx((size_t)nt * g.tnx[0]), hy((size_t)nt * g.tnx[1]), hz((size_t)nt * g.nzc);
    for (size_t m = 0; m < bsq.size(); m++) {
        for (int h = 0; h < 2; h++) {
            const int t = 2 * (int)m + h;
            const double b2 = h == 0 ? bsq[m] : bsq[m] / 2;
            double xf[3];
            for (int d = 0; d < 3; d++) xf[d] = -b2 * std::pow(2 * PI / (g.tnx[d] * g.tdx[d]), 2) * ds / 6.0;
            for (int i = 0; i < g.tnx[0]; i++) {
                const int it = i > g.tnx[0] / 2 ? g.tnx[0] - i : i;
                hx[(size_t)t * g.tnx[0] + i] = std::exp(it * it * xf[0]);
            }
            for (int j = 0; j < g.tnx[1]; j++) {
                const int jt = j > g.tnx[1] / 2 ? g.tnx[1] - j : j;
                hy[(size_t)t * g.tnx[1] + j] = std::exp(jt * jt * xf[1]);
            }
            for (int k = 0; k < g.nzc; k++) hz[(size_t)t * g.nzc + k] = std::exp(k * k * xf[2]);
        }
    }
    auto up = [](DeviceBuffer& b64, DeviceBuffer& b32, const std::vector<double>& h) {
        std::vector<float> f(h.begin(), h.end());
        if (b64.bytes() != sizeof(double) * h.size()) b64.allocate(sizeof(double) * h.size());
        if (b32.bytes() != sizeof(float) * h.size()) b32.allocate(sizeof(float) * h.size());
        FTS_HIP_CHECK(hipMemcpy(b64.get(), h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
        FTS_HIP_CHECK(hipMemcpy(b32.get(), f.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    };
    up(bx64, bx32, hx);
    up(by64, by32, hy);
    up(bz64, bz32, hz);
}

template <>
const double* BondTables::x<double>() const { return bx64.as<double>(); }
template <>
const double* BondTables::y<double>() const { return by64.as<double>(); }
template <>
const double* BondTables::z<double>() const { return bz64.as<double>(); }
template <>
const float* BondTables::x<float>() const { return bx32.as<float>(); }
template <>
const float* BondTables::y<float>() const { return by32.as<float>(); }
template <>
const float* BondTables::z<float>() const { return bz32.as<float>(); }

}  // namespace fts
