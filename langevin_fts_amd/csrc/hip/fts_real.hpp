// Real-space Crank-Nicolson steps for continuous chains (FP64).
//
// Follows CpuSolverReal.cpp:120-451 and FiniteDifference.cpp:6-101 for what one
// step computes: with q_e = exp(-w ds/2) q_in and, per axis b, the tridiagonal
// system T_b = I + B_b (rows l, d, h of FiniteDifference; B_b is half of the
// implicit diffusion operator along b, with the boundary conditions folded in),
//   x_0 = T_0^-1 (I - B_0 - 2 sum_{b>0} B_b) q_e        (first axis, full stencil)
//   x_b = T_b^-1 (x_{b-1} + B_b q_e)                     (each further axis)
//   q_out = exp(-w ds/2) x_last (times the mask).
// The per-axis solves are this file's own: the Thomas factors of every axis are
// computed once on the host (the coefficients do not depend on the line), and a
// periodic axis is solved by the Sherman-Morrison form with gamma = -d[0].
// Lines up to 64 * RT_LMAX points run the tiled sweep (several lanes per line,
// below); longer ones the plain sweep, where one thread owns one line.
#pragma once

#include "fts_common.hpp"
#include "fts_fft.hpp"

namespace fts {

enum { BC_PERIODIC = 0, BC_REFLECTING = 1, BC_ABSORBING = 2 };

// rows of an axis table, each n long, then 2 scalars
enum { RT_L = 0, RT_D, RT_H, RT_INV, RT_C, RT_Z, RT_ROWS };

struct RealAxis {
    int n;         // points along the axis
    int iax;       // internal (3D) axis
    int periodic;
    int off;       // offset of the axis table in a monomer's block
};

struct RealGeom {
    int D;            // real axes, in reference order (x, y, z)
    int tn[3];        // internal 3D sizes (leading unit axes for 1D and 2D)
    long long st[3];  // internal strides
    long long M;
    RealAxis ax[3];
    int tab_stride;   // doubles per monomer
    double half_ds;   // exp(-w ds/2)
};

struct RealItem {
    const double* qin;
    double* qe;     // exp(-w ds/2) q_in
    double* acc;    // the sweeps' intermediate
    double* qout;
    const double* w;
    const double* mask;
    int monomer;
    int need_pre;   // q_e is not yet there: the pre pass computes it from q_in
    int write_qe;   // the last sweep leaves exp(-w ds/2) q_out in q_e for the next step
};

struct RealBatch {
    int n;
    RealItem it[MAX_BATCH];
};
static_assert(sizeof(RealBatch) <= 4000, "real-space batch exceeds the kernel-argument budget");

__global__ void __launch_bounds__(256) real_pre_kernel(RealBatch b, RealGeom g) {
    const RealItem it = b.it[blockIdx.y];
    if (!it.need_pre) return;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < g.M;
         i += (long long)gridDim.x * blockDim.x)
        it.qe[i] = exp(-it.w[i] * g.half_ds) * it.qin[i];
}

// neighbour along an axis: wrapped when periodic, else clamped (its coefficient is 0 then)
__device__ __forceinline__ int rs_prev(int i, int n, int periodic) {
    return i > 0 ? i - 1 : (periodic ? n - 1 : 0);
}
__device__ __forceinline__ int rs_next(int i, int n, int periodic) {
    return i < n - 1 ? i + 1 : (periodic ? 0 : n - 1);
}

// Sweep along real axis a. FIRST: right side from the full explicit stencil of
// q_e; otherwise acc + B_a q_e. LAST: the result goes to q_out times
// exp(-w ds/2) and the mask; otherwise to acc (in place).
template <bool FIRST, bool LAST>
__global__ void __launch_bounds__(256) real_sweep_kernel(RealBatch b, RealGeom g, const double* __restrict__ tab,
                                                          int a) {
    const RealItem it = b.it[blockIdx.y];
    const RealAxis A = g.ax[a];
    const int n = A.n;
    const long long n_lines = g.M / n;
    const long long line = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (line >= n_lines) return;
    // the two internal axes other than this one, the last varying fastest
    const int o1 = A.iax == 0 ? 1 : 0;
    const int o2 = A.iax == 2 ? 1 : 2;
    const int c1 = (int)(line / g.tn[o2]);
    const int c2 = (int)(line % g.tn[o2]);
    const long long base = c1 * g.st[o1] + c2 * g.st[o2];
    const long long sa = g.st[A.iax];
    const double* T = tab + (size_t)it.monomer * g.tab_stride + A.off;
    const double* qe = it.qe;
    double* dst = LAST ? it.qout : it.acc;

    // first sweep: the other axes' part of the stencil is the same for the whole line
    double diag_o = 0.0;
    double cl[2] = {0.0, 0.0}, ch[2] = {0.0, 0.0};
    long long ol[2] = {0, 0}, oh[2] = {0, 0};
    if (FIRST) {
        int k = 0;
        for (int bb = 0; bb < g.D; bb++) {
            if (bb == a) continue;
            const RealAxis B = g.ax[bb];
            const int c = B.iax == o1 ? c1 : c2;
            const double* TB = tab + (size_t)it.monomer * g.tab_stride + B.off;
            diag_o += TB[RT_D * B.n + c];
            cl[k] = TB[RT_L * B.n + c];
            ch[k] = TB[RT_H * B.n + c];
            ol[k] = (rs_prev(c, B.n, B.periodic) - c) * g.st[B.iax];
            oh[k] = (rs_next(c, B.n, B.periodic) - c) * g.st[B.iax];
            k++;
        }
    }

    // forward elimination: y_i = (rhs_i - l_i y_{i-1}) / den_i, stored in dst
    double y = 0.0;
    for (int i = 0; i < n; i++) {
        const long long p = base + i * sa;
        const double q = qe[p];
        const double qm = qe[base + rs_prev(i, n, A.periodic) * sa];
        const double qp = qe[base + rs_next(i, n, A.periodic) * sa];
        const double l = T[RT_L * n + i], d = T[RT_D * n + i], h = T[RT_H * n + i];
        double rhs;
        if (FIRST) {
            double off = 0.0;
            for (int k = 0; k < g.D - 1; k++) off += cl[k] * qe[p + ol[k]] + ch[k] * qe[p + oh[k]];
            rhs = 2.0 * (((double)g.D - 0.5 * d - diag_o) * q - off) - l * qm - h * qp;
        } else {
            rhs = it.acc[p] + (d - 1.0) * q + l * qm + h * qp;
        }
        y = i == 0 ? rhs * T[RT_INV * n] : (rhs - l * y) * T[RT_INV * n + i];
        dst[p] = y;
    }
    // back substitution: x_i = y_i - c_i x_{i+1}
    auto post = [&](long long p, double x) {
        if (LAST) {
            const double E = exp(-it.w[p] * g.half_ds);
            x *= E;
            if (it.mask) x *= it.mask[p];
            // this thread's line is the only reader of these q_e values, and it is done with them
            if (it.write_qe) it.qe[p] = E * x;
        }
        return x;
    };
    const double x_last = y;
    const bool finish_here = !A.periodic;
    {
        const long long p = base + (n - 1) * sa;
        if (finish_here && LAST) dst[p] = post(p, x_last);
    }
    double x = x_last;
    for (int i = n - 2; i >= 0; i--) {
        const long long p = base + i * sa;
        x = dst[p] - T[RT_C * n + i] * x;
        dst[p] = finish_here ? post(p, x) : x;
    }
    if (!finish_here) {
        // x = y - z (y_0 + v y_{n-1}) / (1 + z_0 + v z_{n-1}), both scalars tabulated
        const double v = T[RT_ROWS * n], f = T[RT_ROWS * n + 1];
        const double s = (x + v * x_last) * f;
        for (int i = 0; i < n; i++) {
            const long long p = base + i * sa;
            dst[p] = post(p, dst[p] - T[RT_Z * n + i] * s);
        }
    }
}

// ---- tiled sweeps: P lanes per line, R = 256 / P lines per workgroup
//
// The two recurrences of a line solve, y_i = a_i y_{i-1} + b_i (a_i = -l_i / den_i,
// b_i = rhs_i / den_i) and x_i = g_i x_{i+1} + y_i (g_i = -c_i), are split into P
// chunks of at most RT_LMAX elements, one per lane. Each lane runs its chunk from a
// zero carry, keeping the chunk's affine map (product of the coefficients, end
// value); a shuffle scan over the P lanes composes the maps and hands each lane
// its carry, which the lane then folds into its values. The coefficients are the
// same for every line of an axis, so nothing but the carries travels between lanes.
// The right sides are computed into LDS in coalesced order (rows of the tile along
// the innermost axis), and the results leave LDS the same way.
constexpr int RT_LMAX = 16;
constexpr int RT_THREADS = 256;

struct RealTile {
    int P;        // lanes per line (power of two, at most 64)
    int R;        // lines per workgroup
    int Lc;       // elements per lane (<= RT_LMAX)
    int ktiles;   // outer sweeps: tiles along the innermost axis
};

// INNER: the sweep runs along the innermost (contiguous) axis; the tile is R
// consecutive rows. Otherwise the tile is R consecutive positions of the innermost
// axis at one coordinate of the remaining axis, all along the sweep axis.
template <bool FIRST, bool LAST, bool INNER>
__global__ void __launch_bounds__(RT_THREADS) real_tiled_kernel(RealBatch b, RealGeom g,
                                                                 const double* __restrict__ tab, int a, RealTile t) {
    extern __shared__ double lds[];
    const RealItem it = b.it[blockIdx.y];
    const RealAxis A = g.ax[a];
    const int n = A.n;
    const int R = t.R;
    const double* T = tab + (size_t)it.monomer * g.tab_stride + A.off;
    const double* qe = it.qe;
    double* dst = LAST ? it.qout : it.acc;
    const long long sa = g.st[A.iax];

    // tile origin and valid lines
    long long origin;   // address of element (line 0, position 0)
    long long lstep;    // address step between the tile's lines
    int valid;
    int co = 0, k0 = 0, o = 0;
    if (INNER) {
        const long long rows = g.M / n;
        const long long r0 = (long long)blockIdx.x * R;
        valid = (int)(rows - r0 < R ? rows - r0 : R);
        origin = r0 * n;
        lstep = n;
    } else {
        o = A.iax == 0 ? 1 : 0;   // the internal axis that is neither the sweep axis nor the innermost
        co = blockIdx.x / t.ktiles;
        k0 = (blockIdx.x % t.ktiles) * R;
        valid = g.tn[2] - k0 < R ? g.tn[2] - k0 : R;
        origin = co * g.st[o] + k0;
        lstep = 1;
    }
    const int tid = threadIdx.x;
    const int nel = R * n;

    // the recurrence coefficients of the axis, staged behind the tile:
    // a_i = -l_i / den_i (0 at i = 0), 1 / den_i, g_i = -c_i (0 at i = n-1), z_i
    double* co_a = lds + nel;
    double* co_inv = co_a + n;
    double* co_g = co_inv + n;
    double* co_z = co_g + n;
    for (int i = tid; i < n; i += RT_THREADS) {
        const double inv = T[RT_INV * n + i];
        co_a[i] = i == 0 ? 0.0 : -T[RT_L * n + i] * inv;
        co_inv[i] = inv;
        co_g[i] = i == n - 1 ? 0.0 : -T[RT_C * n + i];
        co_z[i] = T[RT_Z * n + i];
    }

    // ---- right sides into LDS; lines past the grid's edge repeat the last valid
    // line (never written back), which keeps the loop free of branches
#pragma unroll 4
    for (int e = tid; e < nel; e += RT_THREADS) {
        const int r = INNER ? e / n : e % R;
        const int i = INNER ? e % n : e / R;
        const int rc = r < valid ? r : valid - 1;
        const long long line = origin + rc * lstep;
        const long long p = line + i * sa;
        const double q = qe[p];
        const double qm = qe[line + rs_prev(i, n, A.periodic) * sa];
        const double qp = qe[line + rs_next(i, n, A.periodic) * sa];
        const double l = T[RT_L * n + i], d = T[RT_D * n + i], h = T[RT_H * n + i];
        double rhs;
        if (FIRST) {
            double diag_o = 0.0, off = 0.0;
            for (int bb = 0; bb < g.D; bb++) {
                if (bb == a) continue;
                const RealAxis B = g.ax[bb];
                const int c = INNER ? 0 : (B.iax == 2 ? k0 + rc : co);
                const double* TB = tab + (size_t)it.monomer * g.tab_stride + B.off;
                const long long sb = g.st[B.iax];
                diag_o += TB[RT_D * B.n + c];
                off += TB[RT_L * B.n + c] * qe[p + (rs_prev(c, B.n, B.periodic) - c) * sb] +
                       TB[RT_H * B.n + c] * qe[p + (rs_next(c, B.n, B.periodic) - c) * sb];
            }
            rhs = 2.0 * (((double)g.D - 0.5 * d - diag_o) * q - off) - l * qm - h * qp;
        } else {
            rhs = it.acc[p] + (d - 1.0) * q + l * qm + h * qp;
        }
        lds[r * n + i] = rhs;
    }
    __syncthreads();

    // ---- chunked solve: line r = tid / P, chunk c = tid % P
    const int P = t.P, Lc = t.Lc;
    const int r = tid / P, c = tid % P;
    const int s0 = c * Lc;
    const int e0 = s0 + Lc < n ? s0 + Lc : n;
    double v[RT_LMAX];
    double* row = lds + r * n;
    const bool live = r < valid;
    // forward: y_i = a_i y_{i-1} + b_i from a zero carry
    double Am = 1.0, Bm = 0.0;
#pragma unroll
    for (int j = 0; j < RT_LMAX; j++) {
        const int i = s0 + j;
        if (j < Lc && i < e0) {
            const double al = co_a[i];
            const double be = (live ? row[i] : 0.0) * co_inv[i];
            Bm = al * Bm + be;
            Am = al * Am;
            v[j] = Bm;
        }
    }
    for (int dd = 1; dd < P; dd <<= 1) {
        const double Ap = __shfl_up(Am, dd, P), Bp = __shfl_up(Bm, dd, P);
        if (c >= dd) {
            Bm = Am * Bp + Bm;
            Am = Am * Ap;
        }
    }
    double cin = __shfl_up(Bm, 1, P);
    if (c == 0) cin = 0.0;
    {
        double pi = 1.0;
#pragma unroll
        for (int j = 0; j < RT_LMAX; j++) {
            const int i = s0 + j;
            if (j < Lc && i < e0) {
                pi *= co_a[i];
                v[j] += pi * cin;
            }
        }
    }
    // backward: x_i = g_i x_{i+1} + y_i from a zero carry on the right
    Am = 1.0;
    Bm = 0.0;
#pragma unroll
    for (int j = RT_LMAX - 1; j >= 0; j--) {
        const int i = s0 + j;
        if (j < Lc && i < e0) {
            const double ga = co_g[i];
            Bm = ga * Bm + v[j];
            Am = ga * Am;
            v[j] = Bm;
        }
    }
    for (int dd = 1; dd < P; dd <<= 1) {
        const double Ap = __shfl_down(Am, dd, P), Bp = __shfl_down(Bm, dd, P);
        if (c + dd < P) {
            Bm = Am * Bp + Bm;
            Am = Am * Ap;
        }
    }
    cin = __shfl_down(Bm, 1, P);
    if (c == P - 1) cin = 0.0;
    double x_first = 0.0, x_last = 0.0;
    {
        double rho = 1.0;
#pragma unroll
        for (int j = RT_LMAX - 1; j >= 0; j--) {
            const int i = s0 + j;
            if (j < Lc && i < e0) {
                rho *= co_g[i];
                v[j] += rho * cin;
                if (i == 0) x_first = v[j];
                if (i == n - 1) x_last = v[j];
            }
        }
    }
    if (A.periodic) {
        // x = y - z (y_0 + v y_{n-1}) / (1 + z_0 + v z_{n-1})
        const double y0 = __shfl(x_first, 0, P);
        const double yl = __shfl(x_last, (n - 1) / Lc, P);
        const double sc = (y0 + T[RT_ROWS * n] * yl) * T[RT_ROWS * n + 1];
#pragma unroll
        for (int j = 0; j < RT_LMAX; j++) {
            const int i = s0 + j;
            if (j < Lc && i < e0) v[j] -= co_z[i] * sc;
        }
    }
    if (live) {
#pragma unroll
        for (int j = 0; j < RT_LMAX; j++) {
            const int i = s0 + j;
            if (j < Lc && i < e0) row[i] = v[j];
        }
    }
    __syncthreads();

    // ---- results out
#pragma unroll 4
    for (int e = tid; e < nel; e += RT_THREADS) {
        const int rr = INNER ? e / n : e % R;
        const int i = INNER ? e % n : e / R;
        const int rc = rr < valid ? rr : valid - 1;
        const long long p = origin + rc * lstep + i * sa;
        double x = lds[rr * n + i];
        double E = 1.0;
        if (LAST) {
            E = exp(-it.w[p] * g.half_ds);
            x *= E;
            if (it.mask) x *= it.mask[p];
        }
        if (rr < valid) {
            dst[p] = x;
            // the last sweep runs along the innermost axis: the tile's rows are read by
            // this workgroup alone, and all of them were read before the barrier above
            if (LAST && INNER && it.write_qe) it.qe[p] = E * x;
        }
    }
}

}  // namespace fts
