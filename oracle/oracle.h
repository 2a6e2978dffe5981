/*
 * oracle.h -- FP64 CPU restatement of the reference's Gamma / projection path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the GPU path is verified
 * against: only tests/, the smoke test of the repository's entry module and
 * bench.py's CPU-baseline leg may load it, and never as the thing measured or
 * shipped.  The product path (the HIP library
 * under modal-cmb_amd/) does not link, load or call anything in oracle/.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (PetarSuman/Modal-CMB, flat C sources).  Where the reference has a defect
 * the restatement follows the mathematical definition and says so.
 *
 * Pinning (what checks each function independently of this restatement) is
 * listed beside each declaration; see tests/test_oracle.py.
 */
#ifndef MODAL_ORACLE_H
#define MODAL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* tetrapyd_tools.h:31-37 */
typedef struct { int i_bgn, j_bgn, k_bgn, loops; } or_tetrapyd_limits;

/* ---- tetrapyd walks (tetrapyd_tools.c) --------------------------------- */
/* get_ijk_next_XXX  tetrapyd_tools.c:104-142.  Pinned: brute-force count,
 * process.c:65 known answer 334,831,501 at (2,2000). */
int  or_get_ijk_next_XXX(int lmax, int* i, int* j, int* k);
/* check_ijk_min_XXX / check_ijk_max_XXX  tetrapyd_tools.c:144-224 */
void or_check_ijk_min_XXX(int lmin, int* i, int* j, int* k);
void or_check_ijk_max_XXX(int lmax, int* i, int* j, int* k);
/* divide_tasks  tetrapyd_tools.c:595-608 */
void or_divide_tasks(long tasks, int nranks, long* tasks_per_rank);
/* decompose_tetrapyd_XXX  tetrapyd_tools.c:11-102 (count, split, walk to starts) */
void or_decompose_tetrapyd_XXX(or_tetrapyd_limits* block, int rank, int nranks,
                               int nthreads, int lmin, int lmax);
/* number of points of the XXX walk from check_ijk_min(lmin) to its end
 * (tetrapyd_tools.c:32-51) */
long or_count_tetrapyd_XXX(int lmin, int lmax);
/* Dump the walk: writes up to cap (i,j,k) triples, returns count walked. */
long or_walk_XXX(int lmin, int lmax, int* ijk, long cap);
/* Points p0 .. p0+cap-1 of the walk (fewer at its end); returns the count written. */
long or_walk_XXX_window(int lmin, int lmax, long p0, int* ijk, long cap);

/* prim variants, tetrapyd_tools.c:416-592 */
int  or_get_ijk_next_prim(int max, int* i, int* j, int* k);
void or_check_ijk_min_prim(int min, int* i, int* j, int* k);
void or_check_ijk_max_prim(int max, int* i, int* j, int* k);
void or_decompose_tetrapyd_prim(or_tetrapyd_limits* block, int rank, int nranks,
                                int nthreads, int min, int max);
long or_walk_prim(int min, int max, int* ijk, long cap);

/* get_next_ijk_bound  calculate_gamma_3d.c:213-272 (lsize = lmax+1) */
void or_get_next_ijk_bound(int lsize, int i0, int j0, int k0, int* i1, int* j1, int* k1,
                           int final_i, int final_j, int final_k, int* blocksize,
                           int maxblocksize);
/* decompose_gamma_TTT_3D_mpi, host branch (no __MIC__), calculate_gamma_3d.c:92-211.
 * Writes nthreads x {i0,j0,k0,i1,j1,k1}. */
void or_decompose_gamma_TTT_3D(int lsize, int rank, int numranks, double workshare,
                               int nthreads, int* bounds6);

/* ---- weights ------------------------------------------------------------ */
/* permsix  fixed_data.c:4244-4254 */
double or_permsix(int i, int j, int k);
/* calculate_geometric (asymptotic h^2)  calculate_geometric.c:59-92 */
double or_geometric_asym(int l1, int l2, int l3);
/* exact (l1 l2 l3; 0 0 0)^2 and h^2 = (2l1+1)(2l2+1)(2l3+1)/(4 pi) (3j)^2.
 * Pinned: mpmath 40-digit fixtures, sum rule sum_l3 (2l3+1)(3j)^2 = 1. */
double or_threej000_sq(int l1, int l2, int l3);
double or_h2_exact(int l1, int l2, int l3);
/* f(a) = prod_{k<=a} (2k-1)/(2k), a = 0..amax, long-double product rounded. */
void   or_threej_ftable(int amax, double* f);

/* calculate_weight  calculate_volume.c:12-132 (27-neighbour cube fraction
 * times ordered multiplicity) and cell1..cell5 265-310 */
double or_calculate_weight(int min, int max, int i, int j, int k);

/* ---- basis and modes ---------------------------------------------------- */
/* basis_functions_legendre  basis_functions.c:88-107; array is row-major
 * [size][pmax+1].  Pinned: mpmath legendre. */
void or_basis_legendre(double* array, int size, int pmax, double min, double max,
                       const double* values);
/* lexicographic r<=s<=t<R order table (the explicit order file), returns n */
int  or_order_lex(int R, int* order);
/* pijk / plijk_TTT  orthogonalisation.c:279-295, orthogonalisation_late.c:620-644:
 * (1/6) sum over the 6 permutations of q_p1(i) q_p2(j) q_p3(k), q row-major
 * [.][ldq]. */
double or_pijk(const double* q, int ldq, const int* order, int r, int i, int j, int k);

/* Gauss-Legendre nodes/weights on [-1,1], Newton in long double.
 * Pinned: numpy leggauss, mpmath fixtures, exactness on monomials. */
void or_gauss_legendre(int n, double* x, double* w);

/* ---- Gram (Gamma) -------------------------------------------------------- */
/* Direct tetrapyd Gram over T(lmin,lmax) with config weight
 *   w = permsix * h2 * v(l1) v(l2) v(l3),  h2 exact (geom=0) or asymptotic (geom=1)
 * q is [lmax+1][ldq] (row l), v is [lmax+1].  Output full n x n row-major.
 * Restates calculate_orthol_3D's TTT definition (orthogonalisation_late.c:
 * 338-618, without its stale-y1 defect at 608) over all (m,n) pairs at once,
 * as calculate_gamma_TTT_3D_nested does blockwise (calculate_gamma_3d.c:274-440).
 * i-rows [i_lo, i_hi) only when i_hi > 0 (used for timed CPU samples). */
void or_gamma_tet(int lmin, int lmax, const double* q, int ldq, const double* v,
                  int n, const int* order, int geom, int i_lo, int i_hi, double* gamma);
/* Gauss-Legendre-factorised Gram (calculate_orthol, orthogonalisation_late.c:
 * 239-336 with M_ij of calculate_orthoMij_T 99-129, per-leg weight (2l+1) v(l)):
 *   M_ij(x_s) = sum_{l=lmin..lmax} (2l+1) v(l) P_l(x_s) q_i(l) q_j(l)
 *   Gamma_mn  = sum_s w_s/(48 pi) sum_{6 perms rho} prod_a M_{m_a n_rho(a)}(x_s)
 * npt nodes (npt >= (3 lmax + 1)/2 makes it identical to the exact-h2 sum). */
void or_gamma_gl(int lmin, int lmax, const double* q, int ldq, const double* v,
                 int n, const int* order, int npt, double* gamma);

/* ---- polarisation (pol.c) ------------------------------------------------ */
/* XXY walk  tetrapyd_tools.c:226-414: i <= j over [lminX, lmaxX], k over
 * [lminY, lmaxY] with the triangle and even parity, the reference's quirks
 * kept (a new i restarts at k = lminY with no parity bump; the walk ends at
 * i + j + k + 1 >= 2 lmaxX + lmaxY, for an odd lmaxY already at (lmaxX - 1,
 * lmaxX, lmaxY)).  Pinned: brute-force enumeration of
 * the mixed tetrapyd (tests/test_oracle_pol.py). */
int  or_get_ijk_next_XXY(int lminX, int lminY, int lmaxX, int lmaxY, int* i, int* j, int* k);
void or_check_ijk_min_XXY(int lminX, int lminY, int* i, int* j, int* k);
void or_check_ijk_max_XXY(int lmaxX, int lmaxY, int* i, int* j, int* k);
long or_count_tetrapyd_XXY(int lminX, int lminY, int lmaxX, int lmaxY);
long or_walk_XXY(int lminX, int lminY, int lmaxX, int lmaxY, int* ijk, long cap);
void or_decompose_tetrapyd_XXY(or_tetrapyd_limits* block, int rank, int nranks, int nthreads, int lminX,
                               int lminY, int lmaxX, int lmaxY);
/* calculate_orthol's four components (orthogonalisation_late.c:239-336,
 * M^T / M^E of calculate_orthoMij_T / _E 99-157): pol 0 TTT, 1 TTE, 2 TEE,
 * 3 EEE; order in find_perm_late_XXX's slot order (TTE = T, T, E; TEE = T,
 * E, E); q / v of each field are rows by multipole.  Pinned: the direct sum
 * below (TTE, TEE) and the direct tetrapyd Gram on the E field (EEE). */
void or_gamma_gl_pol(int pol, int Tlmin, int Tlmax, const double* qT, int ldqT, const double* vT, int Elmin,
                     int Elmax, const double* qE, int ldqE, const double* vE, int n, const int* order, int npt,
                     double* gamma);
/* Direct sum over the mixed tetrapyd with the exact h^2, multiplicity 2 for
 * unequal pair legs: pol 1 (plijk_TTE over i <= j in T, k in E), pol 2
 * (plijk_TEE over i in T, j <= k in E), orthogonalisation_late.c:646-690. */
void or_gamma_xxy_direct(int pol, int Tlmin, int Tlmax, const double* qT, int ldqT, const double* vT, int Elmin,
                         int Elmax, const double* qE, int ldqE, const double* vE, int n, const int* order,
                         double* gamma);

/* calculate_gamma_2d (spline 0) / calculate_gamma (spline 1) component pol
 * (0 TTT .. 3 EEE) for all (m, n), calculate_gamma.c:48-352, 511-1016: M^T
 * over [Tlmin, Tlmax] with TTw qtT, M^E over [Elmin, Elmax] with EEw (s1 qtE
 * - s2 qtT), late slot b reading the field of the component; spec = 9 rows
 * of lsize_q (C_TT, C_TE, C_EE, N_TT, N_TE, N_EE, b_TT, b_TE, b_EE).
 * Pinned: an x-separable q-tilde reduces every component to a direct sum
 * over the mixed cube with the exact h^2 (tests/test_oracle_pol.py). */
void or_gamma_cross_gl_pol(int pol, int Tlmin, int Tlmax, const double* bT, int ldbT, int Elmin, int Elmax,
                           const double* bE, int ldbE, const int* order_late, int nlate, int np, int lsize_q,
                           const double* qtT, const double* qtE, int nx, const double* xvec, const int* order_prim,
                           int n, const double* spec, double fsky_ratio, double dphi, int spline, int npt,
                           double* gamma);

/* ---- orthonormalisation (ortho_prim_mpi.c:198-284) ---------------------- */
/* norm_i = sqrt(G_ii); C = D^-1 G D^-1; C = L L^T; ortho = symmetrised D L
 * (entry [i][j] = [j][i] = norm[max] L(max,min), update_ortho fixed_data.c:3285);
 * lambda = symmetrised (D L)^-1 (update_lambda fixed_data.c:3576).
 * Returns 0, or -3 if C is not positive definite (the reference aborts). */
int or_orthonormalise(int n, const double* gamma, double* ortho, double* lambda,
                      double* norm);
/* gsl_linalg_cholesky_solve with the symmetrised D L factor
 * (eigenmode_mpi.c:195-219): beta = Gamma^-1 alpha. */
void or_cholesky_solve(int n, const double* ortho, const double* alpha, double* beta);
/* Cyclic Jacobi eigen-decomposition replacing gsl_eigen_symmv + sort
 * descending (ortho_prim_mpi.c:223-224, ortho_late_mpi.c:775-776); evecs[i*n+k]
 * = component i of eigenvector k, largest component positive. */
int or_eigen_sym(int n, const double* A, double* evals, double* evecs);
/* Late-side orthonormalisation with the PCA fallback (ortho_late_mpi.c:752-842). */
int or_ortho_late(int n, const double* G, int pca_flag, double* lambda, double* lambdainv,
                  double* norm, double* evals);

/* ---- shapes and projection ---------------------------------------------- */
/* feature  calculate_shape.c:933-946 ; shape3 = (k1k2k3)^2 shape, 197-216 */
double or_feature(double k1, double k2, double k3, double a1, double a2);
double or_shape3_feature(double k1, double k2, double k3, double a1, double a2);

/* calculate_ortho  orthogonalisation.c:14-86, single-thread semantics (the
 * whole prim tetrapyd, the reference integrates only block 0 when threaded):
 * packed upper triangle, accumulates into ortho_packed. q is [max+1][ldq]. */
void or_calculate_ortho_uniform(int max, double kmax, const double* q, int ldq,
                                int n, const int* order, double* ortho_packed);
/* calculate_eigen  calculate_eigen.c:86-145 for the feature model,
 * accumulates alpha[n]. */
void or_calculate_eigen_uniform(int max, double kmax, const double* q, int ldq, int n,
                                const int* order, double a1, double a2, double* alpha);

/* Gauss-Legendre k-tetrapyd (the build's own quadrature).
 * nodes k_a, weights w_a (a < N), mask: max(k) <= sum of the other two,
 * compared in FP64 as  k3 <= k1 + k2  on the sorted triple.
 * Inner product <f,g> = sum_{cube, mask} w1 w2 w3 f g / (k1+k2+k3). */
/* Gram on that grid; q is [N][ldq] (basis at the nodes) */
void or_gram_kgl(int N, const double* kn, const double* kw, const double* q, int ldq,
                 int n, const int* order, double* gamma);
/* alpha_n(theta) = <sin(omega K + phi), Q_n>; also d alpha / d omega and
 * d alpha / d phi (analytic), any of the outputs may be NULL. */
void or_project_kgl(int N, const double* kn, const double* kw, const double* q, int ldq,
                    int n, const int* order, double omega, double phi,
                    double* alpha, double* dalpha_domega, double* dalpha_dphi);
/* or_project_kgl for nshape pairs theta[b] = (omega, phi), one pass over the
 * points: alpha / d_om / d_ph are [nshape][n] (any may be NULL). */
void or_project_kgl_batch(int N, const double* kn, const double* kw, const double* q, int ldq,
                          int n, const int* order, int nshape, const double* theta,
                          double* alpha, double* d_om, double* d_ph);
/* last index b (of sorted nodes) with kn[b] <= kn[i] + kn[j], per (i,j)
 * (N x N, row-major), -1 if none. */
void or_kgl_limits(int N, const double* kn, int* lim);

/* ---- cross-Gamma against the projected primordial basis ---------------- */
/* init_gamma_3Dint (calculate_gamma_3d.c:23-61, TT part): lw[l] = (2l+1)^-1/3
 * and q3[l][p][x] = (2l+1)^(1/6) / sqrt(C_l + N_l/b_l^2) qt[p][l][x] (l > 1,
 * else 0); qt in the read_qtilde layout [p][l][x] (fixed_data.c:2246-2290). */
void or_qtilde3d(int np, int lsize, int nx, const double* qt, const double* cl, const double* nl,
                 const double* bl, double* q3, double* lw);
/* gamma_3d_host / calculate_gamma_TTT_3D_nested with calculate_sijk_TTT and
 * the trapezium integrate() (gamma_mpi.c:30-152, calculate_gamma_3d.c:
 * 274-440, 783-860): n x n, rows late modes, columns primordial modes, scaled
 * by 6 dphi^2 (nlate x n: the reference's are square); late basis b[l][ldb]
 * by multipole, 0 outside [tlmin, tlmax];
 * points [p0, p1) of the walk of T(2, lsize-1) when p1 > 0. */
void or_gamma_cross_3d(int lsize, int tlmin, int tlmax, const double* b, int ldb, const int* order_late, int nlate,
                       int np, const double* q3, int nx, const double* xvec, const int* order_prim, int n,
                       const double* lw, int geom, double dphi, long p0, long p1, double* gamma);
/* integrate_cubic_spline (calculate_gamma.c:845-934, FMM end conditions; n >= 4:
 * with 2 points its elimination divides 0 by 0, with 3 it reads a 4th) and the natural cubic spline integral of gsl_interp_cspline
 * (calculate_gamma.c:985-998), over [x_0, x_{n-1}]. */
double or_spline_integral_fmm(const double* x, const double* y, int n);
double or_spline_integral_natural(const double* x, const double* y, int n);
/* calculate_gamma_2d (spline 0) / calculate_gamma (spline 1) for all (m, n)
 * after init_gamma_glint + precompute_gammaMij_T (calculate_gamma.c:48-303,
 * 511-555, 772-1016), TT part; scaled by dphi^2 / (8 pi). */
void or_gamma_cross_gl(int tlmin, int tlmax, const double* b, int ldb, const int* order_late, int nlate, int np,
                       int lsize_q, const double* qt, int nx, const double* xvec, const int* order_prim, int n,
                       const double* cl, const double* nl, const double* bl, double dphi, int spline, int npt,
                       double* gamma);

/* wall clock (CLOCK_MONOTONIC) */
double or_wtime(void);
int or_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
