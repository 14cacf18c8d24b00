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
// One thread owns one line; consecutive threads own neighbouring lines, so the
// sweeps along the two outer axes read coalesced rows.
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
};

struct RealBatch {
    int n;
    RealItem it[MAX_BATCH];
};
static_assert(sizeof(RealBatch) <= 4000, "real-space batch exceeds the kernel-argument budget");

__global__ void __launch_bounds__(256) real_pre_kernel(RealBatch b, RealGeom g) {
    const RealItem it = b.it[blockIdx.y];
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
            x *= exp(-it.w[p] * g.half_ds);
            if (it.mask) x *= it.mask[p];
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

}  // namespace fts
