/* CPU oracle for the chain-propagator and concentration path.
 *
 * TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker. The product never links it.
 *
 * FP64 restatement of the reference CPU path:
 *   Pseudo.cpp:17-153                 Boltzmann bond table, weighted Fourier basis
 *   MklFFT3D.cpp:18-19, :55-79        R2C/C2R layout [nx][ny][nz/2+1]; backward / M
 *   CpuSolverPseudo.cpp:89-356        exp(-w ds), steps, half-bond step, segment stress
 *   CpuComputationContinuous.cpp:148-726, CpuComputationDiscrete.cpp:192-996
 *                                     propagators, Q, concentrations, stress
 *   PropagatorComputation.cpp:35-52   stress normalisation by phi/alpha/Q
 *   ComputationBox.cpp:10-308         dv, volume, integral, inner products
 *   CpuAndersonMixing.cpp:76-161, AndersonMixing.cpp:19-47, CircularBuffer.cpp:4-54
 *   FiniteDifference.cpp:6-101, CpuSolverReal.cpp:96-549
 *                                     real-space Crank-Nicolson ADI step (continuous)
 *
 * The FFT is this file's own mixed-radix Cooley-Tukey (recursive decimation in
 * time, exact root table), not MKL; it is pinned against numpy.fft.rfftn and
 * the reference's known-answer tests.
 */
#include "fts_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* ------------------------------------------------------------------ FFT */
typedef struct {
    int n;
    int nf;
    int fac[40];
    cplx* wf; /* exp(-2 pi i j / n) */
    cplx* wb; /* exp(+2 pi i j / n) */
} fft_plan;

static void fft_plan_init(fft_plan* p, int n) {
    p->n = n;
    p->nf = 0;
    int m = n;
    while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
    for (int f = 3; m > 1; f += 2)
        while (m % f == 0) { p->fac[p->nf++] = f; m /= f; }
    p->wf = (cplx*)malloc(sizeof(cplx) * n);
    p->wb = (cplx*)malloc(sizeof(cplx) * n);
    for (int j = 0; j < n; j++) {
        double a = 2.0 * M_PI * (double)j / (double)n;
        p->wf[j] = cos(a) - I * sin(a);
        p->wb[j] = cos(a) + I * sin(a);
    }
}

static void fft_plan_free(fft_plan* p) {
    free(p->wf);
    free(p->wb);
}

/* out[0..n) = DFT_n(in[0], in[is], ...); W: root table of the top length N */
static void fft_rec(const cplx* in, long is, cplx* out, int n, const int* fac, const cplx* W, int N) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    const int p = fac[0], m = n / p, step = N / n;
    for (int r = 0; r < p; r++) fft_rec(in + r * is, is * p, out + r * m, m, fac + 1, W, N);
    cplx t[64];
    cplx* tt = p <= 64 ? t : (cplx*)malloc(sizeof(cplx) * p);
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++) tt[r] = out[r * m + k] * W[(long)r * k * step % N];
        if (p == 2) {
            out[k] = tt[0] + tt[1];
            out[k + m] = tt[0] - tt[1];
        } else if (p == 4) {
            /* W_4^1 = W[N/4] is -i (forward) or +i (backward) */
            cplx j = W[N / 4];
            cplx a0 = tt[0] + tt[2], a1 = tt[0] - tt[2];
            cplx b0 = tt[1] + tt[3], b1 = (tt[1] - tt[3]) * j;
            out[k] = a0 + b0;
            out[k + m] = a1 + b1;
            out[k + 2 * m] = a0 - b0;
            out[k + 3 * m] = a1 - b1;
        } else {
            for (int q = 0; q < p; q++) {
                cplx s = 0;
                for (int r = 0; r < p; r++) s += tt[r] * W[(long)r * q * m * step % N];
                out[k + q * m] = s;
            }
        }
    }
    if (tt != t) free(tt);
}

static void fft_line(const fft_plan* p, const cplx* in, long is, cplx* out, int sign) {
    fft_rec(in, is, out, p->n, p->fac, sign < 0 ? p->wf : p->wb, p->n);
}

/* ------------------------------------------------------------------ grid */
struct orc_grid {
    int dim;
    int nx[3];
    double lx[3];
    double dx[3];
    int tnx[3];
    double tdx[3];
    int M, nzc, Mc;
    double* dv;
    double* mask;
    double volume;
    fft_plan plan[3];
    cplx* work;
};

orc_grid* orc_grid_create(int dim, const int* nx, const double* lx, const double* mask) {
    orc_grid* g = (orc_grid*)calloc(1, sizeof(orc_grid));
    g->dim = dim;
    g->M = 1;
    for (int d = 0; d < dim; d++) {
        g->nx[d] = nx[d];
        g->lx[d] = lx[d];
        g->dx[d] = lx[d] / nx[d];
        g->M *= nx[d];
    }
    /* Pseudo.cpp:31-45: 1D/2D map to leading unit axes with dx = 1 */
    for (int d = 0; d < 3; d++) {
        int src = d - (3 - dim);
        g->tnx[d] = src >= 0 ? nx[src] : 1;
        g->tdx[d] = src >= 0 ? g->dx[src] : 1.0;
    }
    g->nzc = g->tnx[2] / 2 + 1;
    g->Mc = g->tnx[0] * g->tnx[1] * g->nzc;
    g->dv = (double*)malloc(sizeof(double) * g->M);
    if (mask) {
        g->mask = (double*)malloc(sizeof(double) * g->M);
        for (int i = 0; i < g->M; i++) g->mask[i] = fabs(mask[i]) < 1e-7 ? 0.0 : 1.0;
    }
    /* ComputationBox.cpp:75-89 */
    for (int i = 0; i < g->M; i++) {
        g->dv[i] = 1.0;
        for (int d = 0; d < dim; d++) g->dv[i] *= g->dx[d];
        if (g->mask) g->dv[i] *= g->mask[i];
    }
    g->volume = 1.0;
    for (int d = 0; d < dim; d++) g->volume *= g->lx[d];
    for (int d = 0; d < 3; d++) fft_plan_init(&g->plan[d], g->tnx[d]);
    g->work = (cplx*)malloc(sizeof(cplx) * g->Mc);
    return g;
}

void orc_grid_destroy(orc_grid* g) {
    if (!g) return;
    for (int d = 0; d < 3; d++) fft_plan_free(&g->plan[d]);
    free(g->work);
    free(g->dv);
    free(g->mask);
    free(g);
}

/* ComputationBox.cpp:229-262: after set_lx the volume is the sum of dv */
void orc_grid_set_lx(orc_grid* g, const double* lx) {
    for (int d = 0; d < g->dim; d++) {
        g->lx[d] = lx[d];
        g->dx[d] = lx[d] / g->nx[d];
        g->tdx[d + 3 - g->dim] = g->dx[d];
    }
    for (int i = 0; i < g->M; i++) {
        g->dv[i] = 1.0;
        for (int d = 0; d < g->dim; d++) g->dv[i] *= g->dx[d];
        if (g->mask) g->dv[i] *= g->mask[i];
    }
    g->volume = 0.0;
    for (int i = 0; i < g->M; i++) g->volume += g->dv[i];
}

int orc_grid_n(const orc_grid* g) { return g->M; }
int orc_grid_nc(const orc_grid* g) { return g->Mc; }
double orc_grid_volume(const orc_grid* g) { return g->volume; }

double orc_integral(const orc_grid* g, const double* a) {
    double s = 0.0;
    for (int i = 0; i < g->M; i++) s += g->dv[i] * a[i];
    return s;
}

double orc_inner_product(const orc_grid* g, const double* a, const double* b) {
    double s = 0.0;
    for (int i = 0; i < g->M; i++) s += g->dv[i] * a[i] * b[i];
    return s;
}

static double inner_product_inverse_weight(const orc_grid* g, const double* a, const double* b, const double* w) {
    double s = 0.0;
    for (int i = 0; i < g->M; i++) s += g->dv[i] * a[i] * b[i] / w[i];
    return s;
}

/* transform along one axis of the half-complex array: count lines of length n */
static void fft_axis(orc_grid* g, cplx* a, int axis, int sign) {
    const int n0 = g->tnx[0], n1 = g->tnx[1], nc = g->nzc;
    const fft_plan* p = &g->plan[axis];
    const int n = p->n;
    if (n == 1) return;
    const long stride = axis == 0 ? (long)n1 * nc : nc;
    const int lines = axis == 0 ? n1 * nc : n0 * nc;
#pragma omp parallel for schedule(static)
    for (int l = 0; l < lines; l++) {
        cplx buf[n], out[n];
        long base;
        if (axis == 0) base = l;
        else base = (long)(l / nc) * n1 * nc + (l % nc);
        for (int j = 0; j < n; j++) buf[j] = a[base + j * stride];
        fft_line(p, buf, 1, out, sign);
        for (int j = 0; j < n; j++) a[base + j * stride] = out[j];
    }
}

void orc_fft_forward(orc_grid* g, const double* in, double* out_c) {
    const int rows = g->tnx[0] * g->tnx[1], nz = g->tnx[2], nc = g->nzc;
    cplx* a = (cplx*)out_c;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < rows; r++) {
        cplx buf[nz], out[nz];
        for (int k = 0; k < nz; k++) buf[k] = in[(long)r * nz + k];
        fft_line(&g->plan[2], buf, 1, out, -1);
        for (int k = 0; k < nc; k++) a[(long)r * nc + k] = out[k];
    }
    fft_axis(g, a, 1, -1);
    fft_axis(g, a, 0, -1);
}

void orc_fft_backward(orc_grid* g, const double* in_c, double* out) {
    const int rows = g->tnx[0] * g->tnx[1], nz = g->tnx[2], nc = g->nzc;
    cplx* a = g->work;
    memcpy(a, in_c, sizeof(cplx) * g->Mc);
    fft_axis(g, a, 0, +1);
    fft_axis(g, a, 1, +1);
    const double inv = 1.0 / (double)g->M;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < rows; r++) {
        cplx buf[nz], res[nz];
        for (int k = 0; k < nc; k++) buf[k] = a[(long)r * nc + k];
        for (int k = nc; k < nz; k++) buf[k] = conj(a[(long)r * nc + (nz - k)]);
        fft_line(&g->plan[2], buf, 1, res, +1);
        for (int k = 0; k < nz; k++) out[(long)r * nz + k] = creal(res[k]) * inv;
    }
}

/* Pseudo.cpp:17-82 */
void orc_boltz_bond(const orc_grid* g, double bond_length_sq, double ds, double* out) {
    const double PI = 3.14159265358979323846;
    double xf[3];
    for (int d = 0; d < 3; d++) xf[d] = -bond_length_sq * pow(2 * PI / (g->tnx[d] * g->tdx[d]), 2) * ds / 6.0;
    for (int i = 0; i < g->tnx[0]; i++) {
        int it = i > g->tnx[0] / 2 ? g->tnx[0] - i : i;
        for (int j = 0; j < g->tnx[1]; j++) {
            int jt = j > g->tnx[1] / 2 ? g->tnx[1] - j : j;
            for (int k = 0; k < g->nzc; k++) {
                long idx = ((long)i * g->tnx[1] + j) * g->nzc + k;
                out[idx] = exp(it * it * xf[0] + jt * jt * xf[1] + k * k * xf[2]);
            }
        }
    }
}

/* Pseudo.cpp:84-153 */
void orc_fourier_basis(const orc_grid* g, double* bx, double* by, double* bz) {
    const double PI = 3.14159265358979323846;
    double xf[3];
    for (int d = 0; d < 3; d++) xf[d] = pow(2 * PI / (g->tnx[d] * g->tdx[d]), 2);
    for (int i = 0; i < g->tnx[0]; i++) {
        int it = i > g->tnx[0] / 2 ? g->tnx[0] - i : i;
        for (int j = 0; j < g->tnx[1]; j++) {
            int jt = j > g->tnx[1] / 2 ? g->tnx[1] - j : j;
            for (int k = 0; k < g->nzc; k++) {
                long idx = ((long)i * g->tnx[1] + j) * g->nzc + k;
                bx[idx] = it * it * xf[0];
                by[idx] = jt * jt * xf[1];
                bz[idx] = k * k * xf[2];
                if (k != 0 && 2 * k != g->tnx[2]) {
                    bx[idx] *= 2;
                    by[idx] *= 2;
                    bz[idx] *= 2;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ solver */
struct orc_solver {
    orc_grid* g;
    orc_plan pl;
    double** q;   /* per propagator: (max_n+1) slices of M */
    double*** qh; /* per propagator: (max_n+1) pointers, allocated where needed */
    double** E;   /* continuous exp(-w ds/2), discrete exp(-w ds) */
    double** Eh;  /* continuous exp(-w ds/4) */
    double** bb;
    double** bbh;
    double *bx, *by, *bz;
    double* Q;
    double* Qs;
    double** phi;
    double** phis;
    double* dqdl;
    cplx *k1, *k2;
    double *r1, *r2;
    int real;     /* real-space Crank-Nicolson steps instead of pseudo-spectral ones */
    int bc[6];    /* ORC_BC_* per side, reference axis order */
    double* lap;  /* per monomer: dim axes x {l, d, h} x nmax */
    int nmax;
};

static int has_je(const orc_plan* pl, int p, int n) {
    for (int i = 0; i < pl->prop_n_je[p]; i++)
        if (pl->je[pl->prop_je_off[p] + i] == n) return 1;
    return 0;
}

static void update_bond(orc_solver* s) {
    for (int m = 0; m < s->pl.n_monomers; m++) {
        double a = s->pl.segment_length[m];
        /* CpuSolverPseudo.cpp:73-81: b^2 = a*a, half bond uses b^2/2 */
        orc_boltz_bond(s->g, a * a, s->pl.ds, s->bb[m]);
        orc_boltz_bond(s->g, a * a / 2, s->pl.ds, s->bbh[m]);
    }
    orc_fourier_basis(s->g, s->bx, s->by, s->bz);
}

orc_solver* orc_solver_create(orc_grid* g, const orc_plan* plan) {
    orc_solver* s = (orc_solver*)calloc(1, sizeof(orc_solver));
    s->g = g;
    s->pl = *plan;
    const orc_plan* pl = &s->pl;
    const long M = g->M, Mc = g->Mc;
    s->q = (double**)calloc(pl->n_props, sizeof(double*));
    s->qh = (double***)calloc(pl->n_props, sizeof(double**));
    for (int p = 0; p < pl->n_props; p++) {
        int n1 = pl->prop_max_n[p] + 1;
        s->q[p] = (double*)calloc((size_t)n1 * M, sizeof(double));
        s->qh[p] = (double**)calloc(n1, sizeof(double*));
        if (pl->model == ORC_DISCRETE) {
            /* CpuComputationDiscrete.cpp:48-66 */
            if (pl->prop_n_deps[p] > 0) s->qh[p][0] = (double*)calloc(M, sizeof(double));
            for (int n = 1; n < n1; n++)
                if (has_je(pl, p, n)) s->qh[p][n] = (double*)calloc(M, sizeof(double));
        }
    }
    s->E = (double**)calloc(pl->n_monomers, sizeof(double*));
    s->Eh = (double**)calloc(pl->n_monomers, sizeof(double*));
    s->bb = (double**)calloc(pl->n_monomers, sizeof(double*));
    s->bbh = (double**)calloc(pl->n_monomers, sizeof(double*));
    for (int m = 0; m < pl->n_monomers; m++) {
        s->E[m] = (double*)malloc(sizeof(double) * M);
        s->Eh[m] = (double*)malloc(sizeof(double) * M);
        s->bb[m] = (double*)malloc(sizeof(double) * Mc);
        s->bbh[m] = (double*)malloc(sizeof(double) * Mc);
    }
    s->bx = (double*)malloc(sizeof(double) * Mc);
    s->by = (double*)malloc(sizeof(double) * Mc);
    s->bz = (double*)malloc(sizeof(double) * Mc);
    update_bond(s);
    s->Q = (double*)calloc(pl->n_polymers, sizeof(double));
    s->Qs = (double*)calloc(pl->n_solvents > 0 ? pl->n_solvents : 1, sizeof(double));
    s->phi = (double**)calloc(pl->n_blocks, sizeof(double*));
    for (int b = 0; b < pl->n_blocks; b++) s->phi[b] = (double*)calloc(M, sizeof(double));
    s->phis = (double**)calloc(pl->n_solvents > 0 ? pl->n_solvents : 1, sizeof(double*));
    for (int i = 0; i < pl->n_solvents; i++) s->phis[i] = (double*)calloc(M, sizeof(double));
    s->dqdl = (double*)calloc((size_t)pl->n_polymers * 3, sizeof(double));
    s->k1 = (cplx*)malloc(sizeof(cplx) * Mc);
    s->k2 = (cplx*)malloc(sizeof(cplx) * Mc);
    s->r1 = (double*)malloc(sizeof(double) * M);
    s->r2 = (double*)malloc(sizeof(double) * M);
    return s;
}

void orc_solver_destroy(orc_solver* s) {
    if (!s) return;
    const orc_plan* pl = &s->pl;
    for (int p = 0; p < pl->n_props; p++) {
        for (int n = 0; n <= pl->prop_max_n[p]; n++) free(s->qh[p][n]);
        free(s->qh[p]);
        free(s->q[p]);
    }
    free(s->q);
    free(s->qh);
    for (int m = 0; m < pl->n_monomers; m++) {
        free(s->E[m]);
        free(s->Eh[m]);
        free(s->bb[m]);
        free(s->bbh[m]);
    }
    free(s->E);
    free(s->Eh);
    free(s->bb);
    free(s->bbh);
    free(s->bx);
    free(s->by);
    free(s->bz);
    free(s->Q);
    free(s->Qs);
    for (int b = 0; b < pl->n_blocks; b++) free(s->phi[b]);
    free(s->phi);
    for (int i = 0; i < pl->n_solvents; i++) free(s->phis[i]);
    free(s->phis);
    free(s->dqdl);
    free(s->k1);
    free(s->k2);
    free(s->r1);
    free(s->r2);
    free(s->lap);
    free(s);
}

static double* Q_(orc_solver* s, int p, int n) { return s->q[p] + (size_t)n * s->g->M; }

static void kmul(cplx* k, const double* b, long Mc) {
    for (long i = 0; i < Mc; i++) k[i] *= b[i];
}

/* CpuSolverPseudo.cpp:120-188 */
static void step_continuous(orc_solver* s, const double* qin, double* qout, int m) {
    orc_grid* g = s->g;
    const long M = g->M, Mc = g->Mc;
    const double *E = s->E[m], *Eh = s->Eh[m];
    double *q1 = s->r1, *q2 = s->r2;
    for (long i = 0; i < M; i++) q1[i] = E[i] * qin[i];
    orc_fft_forward(g, q1, (double*)s->k1);
    kmul(s->k1, s->bb[m], Mc);
    orc_fft_backward(g, (double*)s->k1, q1);
    for (long i = 0; i < M; i++) q1[i] *= E[i];

    for (long i = 0; i < M; i++) q2[i] = Eh[i] * qin[i];
    orc_fft_forward(g, q2, (double*)s->k2);
    kmul(s->k2, s->bbh[m], Mc);
    orc_fft_backward(g, (double*)s->k2, q2);
    for (long i = 0; i < M; i++) q2[i] *= E[i];
    orc_fft_forward(g, q2, (double*)s->k2);
    kmul(s->k2, s->bbh[m], Mc);
    orc_fft_backward(g, (double*)s->k2, q2);
    for (long i = 0; i < M; i++) q2[i] *= Eh[i];

    for (long i = 0; i < M; i++) qout[i] = (4.0 * q2[i] - q1[i]) / 3.0;
    if (g->mask)
        for (long i = 0; i < M; i++) qout[i] *= g->mask[i];
}

/* CpuSolverPseudo.cpp:189-223 (qin may alias qout) */
static void step_discrete(orc_solver* s, const double* qin, double* qout, int m) {
    orc_grid* g = s->g;
    const long M = g->M;
    orc_fft_forward(g, qin, (double*)s->k1);
    kmul(s->k1, s->bb[m], g->Mc);
    orc_fft_backward(g, (double*)s->k1, qout);
    for (long i = 0; i < M; i++) qout[i] *= s->E[m][i];
    if (g->mask)
        for (long i = 0; i < M; i++) qout[i] *= g->mask[i];
}

/* CpuSolverPseudo.cpp:225-248 */
static void half_bond(orc_solver* s, const double* qin, double* qout, int m) {
    orc_grid* g = s->g;
    orc_fft_forward(g, qin, (double*)s->k1);
    kmul(s->k1, s->bbh[m], g->Mc);
    orc_fft_backward(g, (double*)s->k1, qout);
}


/* ------------------------------------------------------------------ real space */
/* FiniteDifference.cpp:6-101: per axis, l = h = -f/2 and d = 1 + f with
 * f = b^2 ds / (6 dx^2); a reflecting side folds l (or h) into d, an absorbing side
 * drops it. tables: dim blocks of {l, d, h}, each nx[axis] long. */
void orc_laplacian(const orc_grid* g, const int* bc, double bsq, double ds, double* tables) {
    double* t = tables;
    for (int a = 0; a < g->dim; a++) {
        const int n = g->nx[a];
        const double f = bsq * ds / (g->dx[a] * g->dx[a] * 6.0);
        double *l = t, *d = t + n, *h = t + 2 * n;
        for (int i = 0; i < n; i++) {
            l[i] = -0.5 * f;
            d[i] = 1.0 + f;
            h[i] = -0.5 * f;
        }
        if (bc[2 * a] == ORC_BC_REFLECTING) {
            d[0] += l[0];
            l[0] = 0.0;
        } else if (bc[2 * a] == ORC_BC_ABSORBING) {
            l[0] = 0.0;
        }
        if (bc[2 * a + 1] == ORC_BC_REFLECTING) {
            d[n - 1] += h[n - 1];
            h[n - 1] = 0.0;
        } else if (bc[2 * a + 1] == ORC_BC_ABSORBING) {
            h[n - 1] = 0.0;
        }
        t += 3 * n;
    }
}

/* CpuSolverReal.cpp:464-491: Thomas algorithm for rows l[i] x[i-1] + d[i] x[i] + h[i] x[i+1] = rhs[i]
 * (l[0] and h[n-1] unused); x is written with the given stride. */
void orc_tridiagonal(const double* l, const double* d, const double* h, double* x, int stride, const double* rhs,
                     int n) {
    double* c = (double*)malloc(sizeof(double) * (n > 1 ? n : 1));
    double den = d[0];
    x[0] = rhs[0] / den;
    for (int i = 1; i < n; i++) {
        c[i - 1] = h[i - 1] / den;
        den = d[i] - l[i] * c[i - 1];
        x[(long)i * stride] = (rhs[i] - l[i] * x[(long)(i - 1) * stride]) / den;
    }
    for (int i = n - 2; i >= 0; i--) x[(long)i * stride] -= c[i] * x[(long)(i + 1) * stride];
    free(c);
}

/* CpuSolverReal.cpp:493-549: the cyclic system (row 0 couples to x[n-1] through l[0],
 * row n-1 to x[0] through h[n-1]) by the Sherman-Morrison formula with gamma = 1:
 * solve the tridiagonal system with d[0] - 1 and d[n-1] - h[n-1] l[0] for the right
 * side and for u = (1, 0, ..., 0, h[n-1]), then x = y - z (y0 + l0 y_{n-1}) / (1 + z0 + l0 z_{n-1}). */
void orc_tridiagonal_periodic(const double* l, const double* d, const double* h, double* x, int stride,
                              const double* rhs, int n) {
    double* dd = (double*)malloc(sizeof(double) * n);
    double* u = (double*)calloc(n, sizeof(double));
    double* z = (double*)malloc(sizeof(double) * n);
    double* y = (double*)malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) dd[i] = d[i];
    dd[0] -= 1.0;
    dd[n - 1] -= h[n - 1] * l[0];
    u[0] = 1.0;
    u[n - 1] = h[n - 1];
    orc_tridiagonal(l, dd, h, y, 1, rhs, n);
    orc_tridiagonal(l, dd, h, z, 1, u, n);
    const double s = (y[0] + l[0] * y[n - 1]) / (1.0 + z[0] + l[0] * z[n - 1]);
    for (int i = 0; i < n; i++) x[(long)i * stride] = y[i] - z[i] * s;
    free(dd);
    free(u);
    free(z);
    free(y);
}

int orc_solver_set_realspace(orc_solver* s, const int* bc) {
    if (s->pl.model != ORC_CONTINUOUS) return -1;
    const orc_grid* g = s->g;
    s->real = 1;
    for (int i = 0; i < 2 * g->dim; i++) s->bc[i] = bc[i];
    s->nmax = 1;
    for (int a = 0; a < g->dim; a++)
        if (g->nx[a] > s->nmax) s->nmax = g->nx[a];
    free(s->lap);
    s->lap = (double*)calloc((size_t)s->pl.n_monomers * 3 * 3 * s->nmax, sizeof(double));
    return 0;
}

static void update_laplacian_real(orc_solver* s) {
    for (int m = 0; m < s->pl.n_monomers; m++) {
        const double a = s->pl.segment_length[m];
        orc_laplacian(s->g, s->bc, a * a, s->pl.ds, s->lap + (size_t)m * 9 * s->nmax);
    }
}

static void solve_line(const orc_solver* s, int axis, const double* t, double* x, int stride, const double* rhs) {
    const int n = s->g->nx[axis];
    if (s->bc[2 * axis] == ORC_BC_PERIODIC) orc_tridiagonal_periodic(t, t + n, t + 2 * n, x, stride, rhs, n);
    else orc_tridiagonal(t, t + n, t + 2 * n, x, stride, rhs, n);
}

/* neighbour index along an axis: wrap when periodic, else clamp (the coefficient
 * of a clamped neighbour is zero) -- CpuSolverReal.cpp:205-242 */
static int nb(int i, int n, int periodic, int dir) {
    if (periodic) return (i + dir + n) % n;
    i += dir;
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

/* CpuSolverReal.cpp:120-188 with the ADI sweeps of :190-451: q_exp = E q_in; an x solve
 * whose right side holds the full explicit stencil, then one solve per further axis
 * adding back half of that axis' operator; q_out = E x (mask). */
static void step_real(orc_solver* s, const double* qin, double* qout, int m) {
    const orc_grid* g = s->g;
    const int D = g->dim;
    const long M = g->M;
    const double* E = s->E[m];
    double* qe = s->r1;
    double* acc = s->r2;
    for (long i = 0; i < M; i++) qe[i] = E[i] * qin[i];
    int n[3] = {1, 1, 1};
    for (int a = 0; a < D; a++) n[a] = g->nx[a];
    long st[3];
    st[D - 1] = 1;
    for (int a = D - 2; a >= 0; a--) st[a] = st[a + 1] * n[a + 1];
    const double* tab[3];
    for (int a = 0; a < D; a++) tab[a] = s->lap + (size_t)m * 9 * s->nmax + (size_t)3 * (a == 0 ? 0 : (a == 1 ? n[0] : n[0] + n[1]));
    double* rhs = (double*)malloc(sizeof(double) * s->nmax);
    /* first sweep along axis 0, one line per (other coordinates) */
    const long lines0 = M / n[0];
    for (long L = 0; L < lines0; L++) {
        int c[3] = {0, 0, 0};
        long r = L;
        for (int a = D - 1; a >= 1; a--) {
            c[a] = (int)(r % n[a]);
            r /= n[a];
        }
        long base = 0;
        for (int a = 1; a < D; a++) base += c[a] * st[a];
        for (int i = 0; i < n[0]; i++) {
            c[0] = i;
            const long p = base + i * st[0];
            double diag = (double)D - 0.5 * tab[0][n[0] + i];
            double off = 0.0;
            for (int a = 1; a < D; a++) {
                const double* t = tab[a];
                const int per = s->bc[2 * a] == ORC_BC_PERIODIC;
                diag -= t[n[a] + c[a]];
                off += t[c[a]] * qe[p + (nb(c[a], n[a], per, -1) - c[a]) * st[a]] +
                       t[2 * n[a] + c[a]] * qe[p + (nb(c[a], n[a], per, 1) - c[a]) * st[a]];
            }
            const int per0 = s->bc[0] == ORC_BC_PERIODIC;
            rhs[i] = 2.0 * (diag * qe[p] - off) - tab[0][i] * qe[base + nb(i, n[0], per0, -1) * st[0]] -
                     tab[0][2 * n[0] + i] * qe[base + nb(i, n[0], per0, 1) * st[0]];
        }
        solve_line(s, 0, tab[0], (D == 1 ? qout : acc) + base, (int)st[0], rhs);
    }
    /* further sweeps: rhs = previous + (d - 1) q_exp + l q_exp(-) + h q_exp(+) */
    for (int a = 1; a < D; a++) {
        double* dst = a == D - 1 ? qout : acc;
        const long lines = M / n[a];
        const int per = s->bc[2 * a] == ORC_BC_PERIODIC;
        const double* t = tab[a];
        for (long L = 0; L < lines; L++) {
            int c[3] = {0, 0, 0};
            long r = L;
            for (int b = D - 1; b >= 0; b--) {
                if (b == a) continue;
                c[b] = (int)(r % n[b]);
                r /= n[b];
            }
            long base = 0;
            for (int b = 0; b < D; b++)
                if (b != a) base += c[b] * st[b];
            for (int i = 0; i < n[a]; i++) {
                const long p = base + i * st[a];
                rhs[i] = acc[p] + (t[n[a] + i] - 1.0) * qe[p] + t[i] * qe[base + nb(i, n[a], per, -1) * st[a]] +
                         t[2 * n[a] + i] * qe[base + nb(i, n[a], per, 1) * st[a]];
            }
            solve_line(s, a, t, dst + base, (int)st[a], rhs);
        }
    }
    free(rhs);
    for (long i = 0; i < M; i++) qout[i] *= E[i];
    if (g->mask)
        for (long i = 0; i < M; i++) qout[i] *= g->mask[i];
}

int orc_compute_propagators(orc_solver* s, const double* const* w, const double* const* q_init) {
    orc_grid* g = s->g;
    const orc_plan* pl = &s->pl;
    const long M = g->M;
    const double ds = pl->ds;
    if (s->real) update_laplacian_real(s);
    /* CpuSolverPseudo.cpp:89-119 */
    for (int m = 0; m < pl->n_monomers; m++) {
        if (!w[m]) continue;
        for (long i = 0; i < M; i++) {
            if (pl->model == ORC_CONTINUOUS) {
                s->E[m][i] = exp(-w[m][i] * ds * 0.5);
                s->Eh[m][i] = exp(-w[m][i] * ds * 0.25);
            } else {
                s->E[m][i] = exp(-w[m][i] * ds);
            }
        }
    }
    for (int p = 0; p < pl->n_props; p++) {
        const int m = pl->prop_monomer[p], maxn = pl->prop_max_n[p], kind = pl->prop_kind[p];
        const int d0 = pl->prop_dep_off[p], nd = pl->prop_n_deps[p];
        if (pl->model == ORC_CONTINUOUS) {
            /* CpuComputationContinuous.cpp:206-298 */
            double* q0 = Q_(s, p, 0);
            if (kind == ORC_LEAF) {
                for (long i = 0; i < M; i++) q0[i] = 1.0;
            } else if (kind == ORC_LEAF_QINIT) {
                if (!q_init || !q_init[pl->prop_qinit[p]]) return -1;
                memcpy(q0, q_init[pl->prop_qinit[p]], sizeof(double) * M);
            } else if (kind == ORC_AGGREGATED) {
                for (long i = 0; i < M; i++) q0[i] = 0.0;
                for (int d = 0; d < nd; d++) {
                    const double* qd = Q_(s, pl->dep_prop[d0 + d], pl->dep_n[d0 + d]);
                    const int r = pl->dep_rep[d0 + d];
                    for (long i = 0; i < M; i++) q0[i] += qd[i] * r;
                }
            } else {
                for (long i = 0; i < M; i++) q0[i] = 1.0;
                for (int d = 0; d < nd; d++) {
                    const double* qd = Q_(s, pl->dep_prop[d0 + d], pl->dep_n[d0 + d]);
                    for (long i = 0; i < M; i++) q0[i] *= qd[i];
                }
            }
            if (g->mask)
                for (long i = 0; i < M; i++) q0[i] *= g->mask[i];
            for (int n = 0; n < maxn; n++) {
                if (s->real) step_real(s, Q_(s, p, n), Q_(s, p, n + 1), m);
                else step_continuous(s, Q_(s, p, n), Q_(s, p, n + 1), m);
            }
        } else {
            /* CpuComputationDiscrete.cpp:260-541 */
            const double* E = s->E[m];
            double* q1 = Q_(s, p, 1);
            if (kind == ORC_LEAF || kind == ORC_LEAF_QINIT) {
                if (maxn < 1) continue;
                if (kind == ORC_LEAF_QINIT) {
                    if (!q_init || !q_init[pl->prop_qinit[p]]) return -1;
                    for (long i = 0; i < M; i++) q1[i] = q_init[pl->prop_qinit[p]][i] * E[i];
                } else {
                    for (long i = 0; i < M; i++) q1[i] = E[i];
                }
            } else if (kind == ORC_AGGREGATED) {
                for (long i = 0; i < M; i++) q1[i] = 0.0;
                for (int d = 0; d < nd; d++) {
                    const int dp = pl->dep_prop[d0 + d], dn = pl->dep_n[d0 + d], r = pl->dep_rep[d0 + d];
                    const double* qd = dn == 0 ? s->qh[dp][0] : Q_(s, dp, dn);
                    for (long i = 0; i < M; i++) q1[i] += qd[i] * r;
                }
                if (pl->dep_n[d0] == 0) {
                    memcpy(s->qh[p][0], q1, sizeof(double) * M);
                    half_bond(s, q1, q1, m);
                    for (long i = 0; i < M; i++) q1[i] *= E[i];
                } else {
                    step_discrete(s, q1, q1, m);
                }
            } else {
                double* qj = s->qh[p][0];
                for (long i = 0; i < M; i++) qj[i] = 1.0;
                for (int d = 0; d < nd; d++) {
                    const double* hd = s->qh[pl->dep_prop[d0 + d]][pl->dep_n[d0 + d]];
                    for (long i = 0; i < M; i++) qj[i] *= hd[i];
                }
                if (maxn > 0) {
                    half_bond(s, qj, q1, m);
                    for (long i = 0; i < M; i++) q1[i] *= E[i];
                }
            }
            if (maxn == 0) continue;
            if (g->mask)
                for (long i = 0; i < M; i++) q1[i] *= g->mask[i];
            if (s->qh[p][1]) half_bond(s, q1, s->qh[p][1], m);
            for (int n = 1; n < maxn; n++) step_discrete(s, Q_(s, p, n), Q_(s, p, n + 1), m);
            for (int n = 1; n < maxn; n++)
                if (s->qh[p][n + 1]) half_bond(s, Q_(s, p, n + 1), s->qh[p][n + 1], m);
        }
    }
    /* single_partition_segment: first block of each polymer in block order
     * (CpuComputationContinuous.cpp:67-91, CpuComputationDiscrete.cpp:85-115, :563-575) */
    int cur = 0;
    for (int b = 0; b < pl->n_blocks && cur < pl->n_polymers; b++) {
        if (pl->blk_polymer[b] != cur) continue;
        const int nl = pl->blk_n_left[b];
        const int nagg = pl->blk_n_vu[b] / pl->blk_rep[b];
        if (pl->model == ORC_CONTINUOUS) {
            s->Q[cur] = orc_inner_product(g, Q_(s, pl->blk_left[b], nl), Q_(s, pl->blk_right[b], 0)) / nagg /
                        g->volume;
        } else {
            if (nl == 0) continue;
            s->Q[cur] = inner_product_inverse_weight(g, Q_(s, pl->blk_left[b], nl), Q_(s, pl->blk_right[b], 1),
                                                     s->E[pl->blk_monomer[b]]) /
                        nagg / g->volume;
        }
        cur++;
    }
    return 0;
}

/* SimpsonRule.h:10-59 */
static void simpson(int N, double* c) {
    if (N == 0) { c[0] = 1.0; return; }
    if (N == 1) { c[0] = c[1] = 0.5; return; }
    if (N == 3) { c[0] = 3.0 / 8; c[1] = 9.0 / 8; c[2] = 9.0 / 8; c[3] = 3.0 / 8; return; }
    if (N % 2 == 0) {
        c[0] = 1.0 / 3.0;
        for (int n = 1; n <= N - 1; n++) c[n] = n % 2 == 1 ? 4.0 / 3.0 : 2.0 / 3.0;
        c[N] = 1.0 / 3.0;
        return;
    }
    c[0] = 1.0 / 3.0;
    for (int n = 1; n <= N - 4; n++) c[n] = n % 2 == 1 ? 4.0 / 3.0 : 2.0 / 3.0;
    c[N - 3] = 3.0 / 8.0 + 1.0 / 3.0;
    c[N - 2] = 9.0 / 8.0;
    c[N - 1] = 9.0 / 8.0;
    c[N] = 3.0 / 8.0;
}

void orc_simpson(int N, double* out) { simpson(N, out); }

void orc_compute_concentrations(orc_solver* s) {
    orc_grid* g = s->g;
    const orc_plan* pl = &s->pl;
    const long M = g->M;
    for (int b = 0; b < pl->n_blocks; b++) {
        double* phi = s->phi[b];
        const int NR = pl->blk_n_right[b], NL = pl->blk_n_left[b], L = pl->blk_left[b], R = pl->blk_right[b];
        if (NR == 0) {
            for (long i = 0; i < M; i++) phi[i] = 0.0;
            continue;
        }
        if (pl->model == ORC_CONTINUOUS) {
            /* CpuComputationContinuous.cpp:444-465 */
            double* c = (double*)malloc(sizeof(double) * (NR + 1));
            simpson(NR, c);
            const double *a = Q_(s, L, NL), *bq = Q_(s, R, 0);
            for (long i = 0; i < M; i++) phi[i] = c[0] * a[i] * bq[i];
            for (int n = 1; n <= NR; n++) {
                a = Q_(s, L, NL - n);
                bq = Q_(s, R, n);
                for (long i = 0; i < M; i++) phi[i] += c[n] * a[i] * bq[i];
            }
            free(c);
        } else {
            /* CpuComputationDiscrete.cpp:667-690 */
            const double *a = Q_(s, L, NL), *bq = Q_(s, R, 1);
            for (long i = 0; i < M; i++) phi[i] = a[i] * bq[i];
            for (int n = 2; n <= NR; n++) {
                a = Q_(s, L, NL - n + 1);
                bq = Q_(s, R, n);
                for (long i = 0; i < M; i++) phi[i] += a[i] * bq[i];
            }
            const double* E = s->E[pl->blk_monomer[b]];
            for (long i = 0; i < M; i++) phi[i] /= E[i];
        }
        const int p = pl->blk_polymer[b];
        const double norm = pl->ds * pl->pol_vf[p] / pl->pol_alpha[p] / s->Q[p] * pl->blk_rep[b];
        for (long i = 0; i < M; i++) phi[i] *= norm;
    }
    for (int i = 0; i < pl->n_solvents; i++) {
        const double* E = s->E[pl->sol_monomer[i]];
        double* phi = s->phis[i];
        const double vf = pl->sol_vf[i];
        if (pl->model == ORC_CONTINUOUS) {
            /* CpuComputationContinuous.cpp:425-436 */
            s->Qs[i] = orc_inner_product(g, E, E) / g->volume;
            for (long j = 0; j < M; j++) phi[j] = E[j] * E[j] * vf / s->Qs[i];
        } else {
            /* CpuComputationDiscrete.cpp:650-661 */
            s->Qs[i] = orc_integral(g, E) / g->volume;
            for (long j = 0; j < M; j++) phi[j] = E[j] * vf / s->Qs[i];
        }
    }
}

/* CpuSolverPseudo.cpp:250-356: sum_k coeff * basis_d, coeff = b^2 [B] Re(q1_k conj q2_k) */
static void segment_stress(orc_solver* s, const double* q1, const double* q2, double bsq, const double* bond,
                           double* out) {
    orc_grid* g = s->g;
    const long Mc = g->Mc;
    orc_fft_forward(g, q1, (double*)s->k1);
    orc_fft_forward(g, q2, (double*)s->k2);
    const double* basis[3];
    int nd;
    if (g->dim == 3) { basis[0] = s->bx; basis[1] = s->by; basis[2] = s->bz; nd = 3; }
    else if (g->dim == 2) { basis[0] = s->by; basis[1] = s->bz; nd = 2; }
    else { basis[0] = s->bz; nd = 1; }
    double acc[3] = {0, 0, 0};
    for (long i = 0; i < Mc; i++) {
        double coeff;
        if (bond) coeff = bsq * bond[i] * creal(s->k1[i] * conj(s->k2[i]));
        else coeff = bsq * creal(s->k1[i] * conj(s->k2[i]));
        for (int d = 0; d < nd; d++) acc[d] += coeff * basis[d][i];
    }
    for (int d = 0; d < 3; d++) out[d] = acc[d];
}

void orc_compute_stress(orc_solver* s) {
    orc_grid* g = s->g;
    const orc_plan* pl = &s->pl;
    const int DIM = g->dim;
    const long M = g->M;
    double* bdq = (double*)calloc((size_t)pl->n_blocks * 3, sizeof(double));
    for (int b = 0; b < pl->n_blocks; b++) {
        const int NR = pl->blk_n_right[b], NL = pl->blk_n_left[b], L = pl->blk_left[b], R = pl->blk_right[b];
        const int m = pl->blk_monomer[b], rep = pl->blk_rep[b];
        if (NR == 0) continue;
        const double a = pl->segment_length[m];
        double seg[3];
        double acc[3] = {0, 0, 0};
        if (pl->model == ORC_CONTINUOUS) {
            /* CpuComputationContinuous.cpp:671-693 */
            double* c = (double*)malloc(sizeof(double) * (NR + 1));
            simpson(NR, c);
            for (int n = 0; n <= NR; n++) {
                segment_stress(s, Q_(s, L, NL - n), Q_(s, R, n), a * a, NULL, seg);
                for (int d = 0; d < DIM; d++) acc[d] += seg[d] * c[n] * rep;
            }
            free(c);
        } else {
            /* CpuComputationDiscrete.cpp:905-973 */
            for (int n = 0; n <= NR; n++) {
                const double *x1, *x2;
                int half;
                if (n == NL) {
                    if (pl->prop_n_deps[L] == 0) continue;
                    x1 = s->qh[L][0];
                    x2 = Q_(s, R, NR);
                    half = 1;
                } else if (n == 0 && NL == NR) {
                    if (pl->prop_n_deps[R] == 0) continue;
                    x1 = Q_(s, L, NL);
                    x2 = s->qh[R][0];
                    half = 1;
                } else if (n == 0) {
                    continue;
                } else {
                    x1 = Q_(s, L, NL - n);
                    x2 = Q_(s, R, n);
                    half = 0;
                }
                if (half) segment_stress(s, x1, x2, 0.5 * a * a, s->bbh[m], seg);
                else segment_stress(s, x1, x2, a * a, s->bb[m], seg);
                for (int d = 0; d < DIM; d++) acc[d] += seg[d] * rep;
            }
        }
        for (int d = 0; d < 3; d++) bdq[b * 3 + d] = acc[d];
    }
    for (int p = 0; p < pl->n_polymers; p++)
        for (int d = 0; d < 3; d++) s->dqdl[p * 3 + d] = 0.0;
    for (int b = 0; b < pl->n_blocks; b++)
        for (int d = 0; d < DIM; d++) s->dqdl[pl->blk_polymer[b] * 3 + d] += bdq[b * 3 + d];
    for (int p = 0; p < pl->n_polymers; p++)
        for (int d = 0; d < DIM; d++) s->dqdl[p * 3 + d] /= -3.0 * g->lx[d] * M * M / pl->ds;
    free(bdq);
}

double orc_partition(const orc_solver* s, int p) { return s->Q[p]; }
double orc_solvent_partition(const orc_solver* s, int i) { return s->Qs[i]; }
const double* orc_block_phi(const orc_solver* s, int b) { return s->phi[b]; }
const double* orc_solvent_phi(const orc_solver* s, int i) { return s->phis[i]; }
double orc_dq_dl(const orc_solver* s, int p, int d) { return s->dqdl[p * 3 + d]; }

const double* orc_propagator(const orc_solver* s, int prop, int n, int half) {
    if (prop < 0 || prop >= s->pl.n_props || n < 0 || n > s->pl.prop_max_n[prop]) return NULL;
    if (half) return s->qh[prop][n];
    return s->q[prop] + (size_t)n * s->g->M;
}

/* ------------------------------------------------------------------ Anderson mixing */
typedef struct {
    int length, width, start, n_items;
    double* data;
} ring;

static void ring_init(ring* r, int length, int width) {
    r->length = length;
    r->width = width;
    r->start = 0;
    r->n_items = 0;
    r->data = (double*)calloc((size_t)length * width, sizeof(double));
}

/* CircularBuffer.cpp:29-38 */
static void ring_insert(ring* r, const double* v) {
    int i = (r->start + r->n_items) % r->length;
    memcpy(r->data + (size_t)i * r->width, v, sizeof(double) * r->width);
    if (r->n_items == r->length) r->start = (r->start + 1) % r->length;
    r->n_items = r->n_items + 1 < r->length ? r->n_items + 1 : r->length;
}

/* CircularBuffer.cpp:39-43: index 0 is the newest */
static double* ring_get(const ring* r, int n) {
    int i = (r->start + r->n_items - n - 1 + r->length) % r->length;
    return r->data + (size_t)i * r->width;
}

struct orc_am {
    int n_var, max_hist, n_anderson;
    double start_error, mix_min, mix, mix_init;
    ring w_hist, d_hist, dots;
    double *u, *v, *a, *row;
};

orc_am* orc_am_create(int n_var, int max_hist, double start_error, double mix_min, double mix_init) {
    orc_am* a = (orc_am*)calloc(1, sizeof(orc_am));
    a->n_var = n_var;
    a->max_hist = max_hist;
    a->start_error = start_error;
    a->mix_min = mix_min;
    a->mix_init = mix_init;
    ring_init(&a->w_hist, max_hist + 1, n_var);
    ring_init(&a->d_hist, max_hist + 1, n_var);
    ring_init(&a->dots, max_hist + 1, max_hist + 1);
    a->u = (double*)calloc((size_t)max_hist * max_hist + 1, sizeof(double));
    a->v = (double*)calloc(max_hist + 1, sizeof(double));
    a->a = (double*)calloc(max_hist + 1, sizeof(double));
    a->row = (double*)calloc(max_hist + 1, sizeof(double));
    orc_am_reset(a);
    return a;
}

void orc_am_destroy(orc_am* a) {
    if (!a) return;
    free(a->w_hist.data);
    free(a->d_hist.data);
    free(a->dots.data);
    free(a->u);
    free(a->v);
    free(a->a);
    free(a->row);
    free(a);
}

/* CpuAndersonMixing.cpp:50-64 */
void orc_am_reset(orc_am* a) {
    a->mix = a->mix_init;
    a->n_anderson = -1;
    a->w_hist.start = a->w_hist.n_items = 0;
    a->d_hist.start = a->d_hist.n_items = 0;
    a->dots.start = a->dots.n_items = 0;
}

/* AndersonMixing.cpp:19-47: Gaussian elimination without pivoting */
static void find_an(double* u, double* v, double* x, int n, int ld) {
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) {
            double factor = u[j * ld + i] / u[i * ld + i];
            v[j] = v[j] - v[i] * factor;
            for (int k = i + 1; k < n; k++) u[j * ld + k] = u[j * ld + k] - u[i * ld + k] * factor;
        }
    x[n - 1] = v[n - 1] / u[(n - 1) * ld + n - 1];
    for (int i = n - 2; i >= 0; i--) {
        double t = 0.0;
        for (int j = i + 1; j < n; j++) t = t + u[i * ld + j] * x[j];
        x[i] = (v[i] - t) / u[i * ld + i];
    }
}

/* CpuAndersonMixing.cpp:76-161 */
void orc_am_new_fields(orc_am* a, double* w_new, const double* w, const double* d, double old_err, double err) {
    const int n = a->n_var;
    if (err < a->start_error || a->n_anderson >= 0) a->n_anderson = a->n_anderson + 1;
    if (a->n_anderson >= 0) {
        if (a->n_anderson > a->max_hist) a->n_anderson = a->max_hist;
        ring_insert(&a->w_hist, w);
        ring_insert(&a->d_hist, d);
        for (int i = 0; i <= a->n_anderson; i++) {
            const double* h = ring_get(&a->d_hist, i);
            double s = 0.0;
            for (int j = 0; j < n; j++) s += d[j] * h[j];
            a->row[i] = s;
        }
        ring_insert(&a->dots, a->row);
    }
    if (a->n_anderson <= 0) {
        if (old_err < err) a->mix = a->mix * 0.7 > a->mix_min ? a->mix * 0.7 : a->mix_min;
        else a->mix = a->mix * 1.01;
        for (int i = 0; i < n; i++) w_new[i] = w[i] + a->mix * d[i];
        return;
    }
    const int na = a->n_anderson, ld = a->max_hist;
    const double* d0 = ring_get(&a->dots, 0);
    for (int i = 0; i < na; i++) {
        a->v[i] = d0[0] - d0[i + 1];
        for (int j = 0; j < na; j++) {
            int r = (i + 1) < (j + 1) ? (i + 1) : (j + 1);
            int c = i - j < 0 ? j - i : i - j;
            a->u[i * ld + j] = d0[0] - d0[i + 1] - d0[j + 1] + ring_get(&a->dots, r)[c];
        }
    }
    find_an(a->u, a->v, a->a, na, ld);
    const double* w1 = ring_get(&a->w_hist, 0);
    const double* e1 = ring_get(&a->d_hist, 0);
    for (int i = 0; i < n; i++) w_new[i] = w1[i] + e1[i];
    for (int k = 0; k < na; k++) {
        const double* w2 = ring_get(&a->w_hist, k + 1);
        const double* e2 = ring_get(&a->d_hist, k + 1);
        for (int j = 0; j < n; j++) w_new[j] += a->a[k] * (w2[j] + e2[j] - w1[j] - e1[j]);
    }
}
