// CPU block emulator of the generic FFT pass kernels of fts_fft.hpp (slab_kernel,
// xpass_kernel): the kernels' own source, compiled for the host against the stand-in
// hip/hip_runtime.h beside this file. Each emulated GPU thread runs the kernel in its
// own std::thread; __syncthreads is a barrier that reports a deadlock (a thread
// waiting on a barrier the others never reach) instead of hanging; threads that
// return leave the barrier's count, as finished wavefronts do. The workgroup's LDS is
// a global array of exactly the kernel's dynamic LDS size (LDS_BYTES), filled with
// NaN before each block, so that with -fsanitize=address every out-of-range LDS index
// is reported and a read of never-written LDS poisons the result.
//
//   fft_block_emu <slab|xpass> <fwd|inv> <precision 4|8> nx ny nz <in> <out>
//
// slab fwd: real fields -> half spectra after the y and z transforms (slab FWD mode);
// xpass fwd: x forward transform in place (table -1); xpass inv: x inverse (table -2);
// slab inv: half spectra -> real fields (slab INV1 mode, 1/M included). The launch
// geometry is Grid::setup_fft's for the grid (restated below); it prints the geometry
// and the barrier count of every thread.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "fts_fft.hpp"

#ifndef LDS_BYTES
#error "LDS_BYTES: the kernel's dynamic LDS size"
#endif

thread_local dim3 threadIdx;
thread_local dim3 blockIdx;
dim3 blockDim;
dim3 gridDim;

namespace fts {
alignas(16) unsigned char smem_raw[LDS_BYTES];
}

namespace {

struct Barrier {
    std::mutex m;
    std::condition_variable cv;
    int expected = 0, waiting = 0;
    long generation = 0;
    bool dead = false;
} bar;
thread_local long my_barriers = 0;

void release_locked() {
    bar.waiting = 0;
    bar.generation++;
    bar.cv.notify_all();
}

}  // namespace

void emu_syncthreads() {
    std::unique_lock<std::mutex> lk(bar.m);
    my_barriers++;
    const long gen = bar.generation;
    if (++bar.waiting == bar.expected) {
        release_locked();
        return;
    }
    if (!bar.cv.wait_for(lk, std::chrono::seconds(10), [&] { return bar.generation != gen; })) {
        bar.dead = true;
        std::fprintf(stderr, "DEADLOCK: thread %u of block (%u, %u) waits at barrier %ld; %d of %d arrived\n",
                     threadIdx.x, blockIdx.x, blockIdx.y, my_barriers, bar.waiting, bar.expected);
        std::fflush(stderr);
        std::_Exit(3);
    }
}

namespace {

void thread_exit() {
    std::lock_guard<std::mutex> lk(bar.m);
    bar.expected--;
    if (bar.waiting > 0 && bar.waiting == bar.expected) release_locked();
}

template <typename F>
void run_grid(dim3 grid, dim3 block, size_t lds, F&& kernel) {
    if (lds != LDS_BYTES) {
        std::fprintf(stderr, "LDS size %zu differs from the compiled %d\n", lds, LDS_BYTES);
        std::exit(2);
    }
    gridDim = grid;
    blockDim = block;
    for (unsigned by = 0; by < grid.y; by++)
        for (unsigned bx = 0; bx < grid.x; bx++) {
            const float nan = std::nanf("");
            for (size_t i = 0; i < LDS_BYTES / sizeof(float); i++) std::memcpy(fts::smem_raw + 4 * i, &nan, 4);
            bar.expected = (int)block.x;
            bar.waiting = 0;
            std::vector<long> counts(block.x);
            std::vector<std::thread> th;
            for (unsigned t = 0; t < block.x; t++)
                th.emplace_back([&, t] {
                    threadIdx = dim3{t, 0, 0};
                    blockIdx = dim3{bx, by, 0};
                    my_barriers = 0;
                    kernel();
                    counts[t] = my_barriers;
                    thread_exit();
                });
            for (auto& x : th) x.join();
            for (unsigned t = 1; t < block.x; t++)
                if (counts[t] != counts[0]) {
                    std::fprintf(stderr, "BARRIER MISMATCH in block (%u, %u): thread 0 %ld, thread %u %ld\n", bx, by,
                                 counts[0], t, counts[t]);
                    std::exit(4);
                }
            if (bx == 0 && by == 0) std::printf("barriers per thread %ld\n", counts[0]);
        }
}

// Grid::setup_fft's choice of workgroup size: about 8 elements per thread, whole waves
int choose_threads(int work) {
    int bd = fts::ceil_div(fts::ceil_div(work, 8), 64) * 64;
    return std::max(64, std::min(1024, bd));
}

template <typename T>
std::vector<typename fts::vec2<T>::type> twiddles(int L) {
    std::vector<typename fts::vec2<T>::type> h(L);
    for (int m = 0; m < L; m++) {
        const double a = 2.0 * M_PI * (double)m / (double)L;
        h[m] = {(T)std::cos(a), (T)-std::sin(a)};
    }
    return h;
}

template <typename T>
int run(const char* pass, const char* dir, const int* n, const char* in_path, const char* out_path) {
    using T2 = typename fts::vec2<T>::type;
    const int nx = n[0], ny = n[1], nz = n[2], nzc = nz / 2 + 1;
    const size_t M = (size_t)nx * ny * nz, Mc = (size_t)nx * ny * nzc;
    const fts::AxisPlan ax = fts::plan_axis(nx), ay = fts::plan_axis(ny), az = fts::plan_axis(nz);
    const bool big = fts::plan_is_big(ax) || fts::plan_is_big(ay) || fts::plan_is_big(az);
    const bool fwd = std::strcmp(dir, "fwd") == 0;
    const bool slab = std::strcmp(pass, "slab") == 0;
    const size_t n_in = slab && fwd ? M : 2 * Mc;
    const size_t n_out = slab && !fwd ? M : 2 * Mc;
    std::vector<T> in(n_in), out(n_out);
    FILE* f = std::fopen(in_path, "rb");
    if (!f || std::fread(in.data(), sizeof(T), n_in, f) != n_in) return 2;
    std::fclose(f);

    if (slab) {
        fts::SlabGeom g{};
        g.ny = ny;
        g.nz = nz;
        g.nzc = nzc;
        g.P = nzc;
        g.rows = 2 * ((ny + 1) / 2);
        g.ay = ay;
        g.az = az;
        g.inv_m = 1.0 / (double)M;
        const int threads = choose_threads(ny * nz);
        const int per_thread = fts::ceil_div(ny * nz, threads);
        const int maxr = (per_thread <= 8 && threads <= 512) ? 8 : 16;
        g.ay.group = fts::group_lines(g.ay, threads);
        g.az.group = fts::group_lines(g.az, threads);
        const size_t lds = (size_t)(g.rows * g.P + g.ny + g.nz) * sizeof(T2);
        std::printf("slab: threads %d, maxr %d, big %d, y group %d, z group %d, lds %zu\n", threads, maxr, big,
                    g.ay.group, g.az.group, lds);
        auto twy = twiddles<T>(ny), twz = twiddles<T>(nz);
        fts::SlabBatch<T> b{};
        b.n = 1;
        if (fwd) {
            b.it[0].qin = in.data();
            b.it[0].c0 = reinterpret_cast<T2*>(out.data());
            b.it[0].n_out = 1;
        } else {
            b.it[0].c0 = reinterpret_cast<T2*>(in.data());
            b.it[0].qout = out.data();
        }
        const dim3 grid{(unsigned)nx, 1, 1}, block{(unsigned)threads, 1, 1};
        auto go = [&](auto kern) { run_grid(grid, block, lds, [&] { kern(g, b, twy.data(), twz.data()); }); };
        if (fwd) {
            if (maxr == 8) big ? go(fts::slab_kernel<T, fts::SLAB_FWD, 8, true>) : go(fts::slab_kernel<T, fts::SLAB_FWD, 8, false>);
            else big ? go(fts::slab_kernel<T, fts::SLAB_FWD, 16, true>) : go(fts::slab_kernel<T, fts::SLAB_FWD, 16, false>);
        } else {
            if (maxr == 8) big ? go(fts::slab_kernel<T, fts::SLAB_INV1, 8, true>) : go(fts::slab_kernel<T, fts::SLAB_INV1, 8, false>);
            else big ? go(fts::slab_kernel<T, fts::SLAB_INV1, 16, true>) : go(fts::slab_kernel<T, fts::SLAB_INV1, 16, false>);
        }
    } else {
        fts::XGeom g{};
        g.nx = nx;
        g.ny = ny;
        g.nzc = nzc;
        int C = nzc, tiles = 1;
        while (ny * tiles < 128 && C > 8) {
            C = fts::ceil_div(C, 2);
            tiles = fts::ceil_div(nzc, C);
        }
        while ((size_t)(nx * C + nx) * sizeof(double2) > 64 * 1024 && C > 1) {
            C = fts::ceil_div(C, 2);
            tiles = fts::ceil_div(nzc, C);
        }
        g.C = C;
        g.n_tiles = tiles;
        g.P = C;
        g.ax = ax;
        const int threads = choose_threads(nx * C);
        g.ax.group = fts::group_lines(g.ax, threads);
        const size_t lds = (size_t)(nx * g.P + nx) * sizeof(T2);
        std::printf("xpass: threads %d, tile %d x %d, big %d, x group %d, lds %zu\n", threads, C, tiles, big,
                    g.ax.group, lds);
        out = in;   // the pass works in place
        auto twx = twiddles<T>(nx);
        fts::XBatch<T> b{};
        b.n = 1;
        b.it[0] = fts::XItem<T>{reinterpret_cast<T2*>(out.data()), fwd ? -1 : -2};
        const dim3 grid{(unsigned)(ny * tiles), 1, 1}, block{(unsigned)threads, 1, 1};
        auto go = [&](auto kern) {
            run_grid(grid, block, lds, [&] { kern(g, b, twx.data(), (const T*)nullptr, (const T*)nullptr, (const T*)nullptr); });
        };
        big ? go(fts::xpass_kernel<T, true>) : go(fts::xpass_kernel<T, false>);
    }
    f = std::fopen(out_path, "wb");
    if (!f || std::fwrite(out.data(), sizeof(T), n_out, f) != n_out) return 2;
    std::fclose(f);
    std::printf("ok\n");
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 9) {
        std::fprintf(stderr, "usage: %s <slab|xpass> <fwd|inv> <4|8> nx ny nz in out\n", argv[0]);
        return 2;
    }
    const int n[3] = {std::atoi(argv[4]), std::atoi(argv[5]), std::atoi(argv[6])};
    return std::atoi(argv[3]) == 4 ? run<float>(argv[1], argv[2], n, argv[7], argv[8])
                                   : run<double>(argv[1], argv[2], n, argv[7], argv[8]);
}
