/* CPU oracle for the chain-propagator and concentration path (test infrastructure).
 *
 * A plain-C FP64 restatement of the reference CPU path, used only as the checker
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg. It is never
 * linked into the product. Each routine cites the reference file:line it follows.
 *
 * The computation consumes a "plan": the deduplicated propagators and computation
 * blocks that PropagatorAnalyzer produces (PropagatorAnalyzer.cpp:142-269), given as
 * flat integer arrays. Propagators are evaluated in plan order, which is the
 * ComparePropagatorKey order (height ascending), so every dependency is ready.
 */
#ifndef FTS_ORACLE_H
#define FTS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid orc_grid;
typedef struct orc_solver orc_solver;
typedef struct orc_am orc_am;

enum { ORC_CONTINUOUS = 0, ORC_DISCRETE = 1 };
enum { ORC_LEAF = 0, ORC_LEAF_QINIT = 1, ORC_JUNCTION = 2, ORC_AGGREGATED = 3 };

/* grid: dim 1..3, nx/lx of length dim, mask NULL or M entries of 0/1 */
orc_grid* orc_grid_create(int dim, const int* nx, const double* lx, const double* mask);
void orc_grid_destroy(orc_grid* g);
void orc_grid_set_lx(orc_grid* g, const double* lx);
int orc_grid_n(const orc_grid* g);
int orc_grid_nc(const orc_grid* g);
double orc_grid_volume(const orc_grid* g);
double orc_integral(const orc_grid* g, const double* a);
double orc_inner_product(const orc_grid* g, const double* a, const double* b);

/* FP64 real<->half-complex 3D transforms on the grid (1D/2D mapped to 3D).
 * forward unnormalised, backward divides by M. complex = interleaved re,im. */
void orc_fft_forward(orc_grid* g, const double* in, double* out_c);
void orc_fft_backward(orc_grid* g, const double* in_c, double* out);

/* Boltzmann bond table exp(-b^2 ds/6 |k|^2) on the half spectrum; basis_d = (2 pi k_d/L_d)^2
 * with the half-spectrum weight (x2 for 0 < kz < nz/2). */
void orc_boltz_bond(const orc_grid* g, double bond_length_sq, double ds, double* out);
void orc_fourier_basis(const orc_grid* g, double* bx, double* by, double* bz);

typedef struct {
    int model;
    double ds;
    int n_monomers;
    const double* segment_length; /* a_m; b^2 = a_m^2 */
    int n_props;
    const int* prop_monomer;
    const int* prop_max_n;
    const int* prop_kind;
    const int* prop_qinit;
    const int* prop_dep_off;
    const int* prop_n_deps;
    const int* dep_prop;
    const int* dep_n;
    const int* dep_rep;
    const int* prop_je_off;
    const int* prop_n_je;
    const int* je;
    int n_blocks;
    const int* blk_polymer;
    const int* blk_left;
    const int* blk_right;
    const int* blk_n_left;
    const int* blk_n_right;
    const int* blk_rep;
    const int* blk_monomer;
    const int* blk_n_vu;
    int n_polymers;
    const double* pol_vf;
    const double* pol_alpha;
    int n_solvents;
    const double* sol_vf;
    const int* sol_monomer;
} orc_plan;

orc_solver* orc_solver_create(orc_grid* g, const orc_plan* plan);
void orc_solver_destroy(orc_solver* s);
/* w: n_monomers pointers of M doubles; q_init: pointers indexed by prop_qinit (may be NULL) */
int orc_compute_propagators(orc_solver* s, const double* const* w, const double* const* q_init);
void orc_compute_concentrations(orc_solver* s);
void orc_compute_stress(orc_solver* s);
double orc_partition(const orc_solver* s, int p);
double orc_solvent_partition(const orc_solver* s, int i);
const double* orc_block_phi(const orc_solver* s, int b);
const double* orc_solvent_phi(const orc_solver* s, int i);
/* dq_dl[p][d], d < dim */
double orc_dq_dl(const orc_solver* s, int p, int d);
/* q[prop][n]; half = 1 selects the half-bond array (discrete) */
const double* orc_propagator(const orc_solver* s, int prop, int n, int half);

/* Real-space Crank-Nicolson ADI (continuous chains): boundary condition per side,
 * reference axis order (x low, x high, y low, ...). */
enum { ORC_BC_PERIODIC = 0, ORC_BC_REFLECTING = 1, ORC_BC_ABSORBING = 2 };
/* switch the solver to real-space steps (CpuComputationContinuous.cpp:25-26); -1 for discrete chains */
int orc_solver_set_realspace(orc_solver* s, const int* bc);
/* FiniteDifference::get_laplacian_matrix: tables = dim blocks of {l, d, h}, nx[axis] each */
void orc_laplacian(const orc_grid* g, const int* bc, double bond_length_sq, double ds, double* tables);
/* CpuSolverReal::tridiagonal / tridiagonal_periodic: solve for x (written with stride) */
void orc_tridiagonal(const double* l, const double* d, const double* h, double* x, int stride, const double* rhs,
                     int n);
void orc_tridiagonal_periodic(const double* l, const double* d, const double* h, double* x, int stride,
                              const double* rhs, int n);
/* SimpsonRule.h:10-59 coefficients for N intervals (N+1 values) */
void orc_simpson(int N, double* out);

/* Anderson mixing, CpuAndersonMixing.cpp:76-161 with AndersonMixing.cpp:19-47 */
orc_am* orc_am_create(int n_var, int max_hist, double start_error, double mix_min, double mix_init);
void orc_am_destroy(orc_am* a);
void orc_am_reset(orc_am* a);
void orc_am_new_fields(orc_am* a, double* w_new, const double* w, const double* d, double old_err, double err);

#ifdef __cplusplus
}
#endif
#endif
