/*
 * modal_cmb.h -- C ABI of the MI355X (gfx950) Gamma / shape-projection library.
 *
 * Plain C types only (pointers, sizes, status codes).  Two layers:
 *
 *  1. An explicit-state API over an opaque context (one GPU, one stream).
 *     Every call returns an int status (MCMB_OK == 0, negative on error) and
 *     never exits; mcmb_last_error() gives the message of the calling
 *     thread's last failure.
 *  2. The reference's own entry points (same names, argument meaning and
 *     array layouts as PetarSuman/Modal-CMB), implemented as shims over a
 *     process-global context bound with mcmb_legacy_bind().  Each declaration
 *     cites the reference prototype it replaces.
 *
 * Array layouts follow the reference: row-major FP64, n x n Gram matrices,
 * basis tables [point][pmax+1], mode orders [mode][3].
 */
#ifndef MODAL_CMB_H
#define MODAL_CMB_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define MCMB_OK            0
#define MCMB_ERR_ARG      -1   /* bad argument or size                          */
#define MCMB_ERR_HIP      -2   /* HIP runtime failure (message has the detail)  */
#define MCMB_ERR_NOT_SPD  -3   /* Cholesky pivot <= 0: Gram not positive definite */
#define MCMB_ERR_ORDER    -4   /* bad mode-order table                          */
#define MCMB_ERR_IO       -5   /* file missing, short or unwritable             */
#define MCMB_ERR_NOMEM    -6   /* device or host allocation failed              */
#define MCMB_ERR_STATE    -7   /* required setup call not made                  */

/* precision of the tetrapyd Gram contraction */
#define MCMB_FP64   0   /* FP64 MFMA, FP64 accumulation                          */
#define MCMB_MIXED  1   /* BF16 hi+lo split operands on BF16 MFMA (3 products),
                           FP32 chunk accumulation, FP64 cross-chunk sums, FP64
                           diagonal lo*lo correction                              */

#define MCMB_FP32   2   /* projection only: FP32 MFMA k-sum, FP64 mode sums    */

/* x integration of the cross-Gamma */
#define MCMB_SPLINE_FMM        0  /* integrate_cubic_spline (calculate_gamma_2d)    */
#define MCMB_SPLINE_NATURAL    1  /* gsl_interp_cspline (calculate_gamma)           */
#define MCMB_SPLINE_TRAPEZIUM  2  /* integrate() with TRAPEZIUM (gamma_3d_host)     */

/* geometric factor h^2 */
#define MCMB_H2_EXACT      0  /* (2l1+1)(2l2+1)(2l3+1)/(4pi) (l1 l2 l3;0 0 0)^2   */
#define MCMB_H2_ASYMPTOTIC 1  /* calculate_geometric.c:59-92                     */

/* polarisation components (calculate_orthol / calculate_gamma result[0..3]) */
#define MCMB_POL_TTT 0
#define MCMB_POL_TTE 1
#define MCMB_POL_TEE 2
#define MCMB_POL_EEE 3

typedef struct mcmb_ctx mcmb_ctx;

/* tetrapyd_tools.h:31-37 */
typedef struct { int i_bgn; int j_bgn; int k_bgn; int loops; } tetrapyd_limits;
/* calculate_gamma_3d.h:12-29: per-thread data of the 3D Gamma; the library
 * reads and writes only the walk bounds i0..k1 (bar is the reference's
 * barrier_t pointer, kept for the layout). */
typedef struct {
    int i0; int j0; int k0; int i1; int j1; int k1;
    int xsize;
    double* mvec; double* xvec; double* yvec; double* xdiff; double* ixdiff;
    double* intgrlvec; double* plijkz;
    void* bar;
} thread_g3d_data_t;
/* global.h:32-36 */
typedef struct { double a1; double a2; double a3; } shape_params;

const char* mcmb_last_error(void);
const char* mcmb_status_string(int status);

/* Create a context on HIP device `device` (own non-blocking stream). */
int mcmb_create(mcmb_ctx** ctx, int device);
int mcmb_destroy(mcmb_ctx* ctx);
/* Block until all work queued on the context's stream has finished. */
int mcmb_sync(mcmb_ctx* ctx);

/* ---- late-time (l-space) Gram --------------------------------------------- */
/* Load the 1D basis q[l][0..ldq) for l = 0..lmax (row l is multipole l; only
 * l in [lmin,lmax] is read), per-leg weight v[l] (= 1/C_l for the config
 * weight h^2/(C_l1 C_l2 C_l3)), and the mode order table order[n][3]
 * (r,s,t < ldq, ldq <= 32).  Replaces the state create_basis_late / set_terms_late
 * leave behind (fixed_data.c:272-1986).  Tables are copied to the device. */
int mcmb_lspace_set(mcmb_ctx* ctx, int lmin, int lmax, int ldq, const double* q,
                    const double* v, int nmodes, const int* order, int h2_mode);

/* Gamma_mn = sum_{l1<=l2<=l3 in T(lmin,lmax)} permsix h^2 v1 v2 v3 Q_m Q_n
 * (the tetrapyd contraction of calculate_orthol_3D, orthogonalisation_late.c:
 * 338-618, and of calculate_gamma_TTT_3D_nested, calculate_gamma_3d.c:274-440,
 * without the 6 dphi^2 scaling).  Writes the full symmetric n x n matrix to
 * gamma_host (row-major) if non-NULL, and/or to the device buffer gamma_dev
 * (n x n doubles) if non-NULL. */
int mcmb_gamma_tet(mcmb_ctx* ctx, int precision, double* gamma_host, double* gamma_dev);

/* Partial Gram over rank `rank` of `nranks` contiguous shares of the walk
 * (balanced by point chunks), the decomposition decompose_tetrapyd_XXX makes
 * for MPI ranks (tetrapyd_tools.c:11-102); the shares of all ranks sum to
 * mcmb_gamma_tet.  *work_fraction (if non-NULL) = share of the chunks. */
int mcmb_gamma_tet_slice(mcmb_ctx* ctx, int precision, int rank, int nranks, double* gamma_host,
                         double* gamma_dev, double* work_fraction);

/* Partial Gram over the points with l1 in [l1_lo, l1_hi) (a contiguous range of
 * the walk: the i-rows the loop nest of calculate_orthol_3D visits,
 * orthogonalisation_late.c:548-615); ranges that tile [lmin, lmax] sum to
 * mcmb_gamma_tet.  No reference counterpart: checked i-row samples. */
int mcmb_gamma_tet_lrange(mcmb_ctx* ctx, int precision, int l1_lo, int l1_hi, double* gamma_host,
                          double* gamma_dev);

/* The same Gram through the Gauss-Legendre factorisation (calculate_orthol,
 * orthogonalisation_late.c:239-336) with npt nodes (npt <= 0: 3 lmax/2 + 1). */
int mcmb_gamma_gl(mcmb_ctx* ctx, int npt, double* gamma_host, double* gamma_dev);

/* E-mode field of the late basis for the polarised Gram (create_basis_late's
 * E columns, fixed_data.c:1923-1950; the TTE / TEE / EEE tables of
 * set_terms_late, fixed_data.c:272-417): basis qE[l][ldq] for l = 0..E_lmax
 * (rows below E_lmin unused), per-leg weight vE[l], and nmodes rows per
 * component table in find_perm_late_XXX's slot order: order_TTE (T, T, E),
 * order_TEE (T, E, E) -- the reference's TEE file lists (E, E, T) and
 * set_terms_late reorders it --, order_EEE (E, E, E).  T indices refer to
 * the basis of mcmb_lspace_set. */
int mcmb_lspace_set_E(mcmb_ctx* ctx, int E_lmin, int E_lmax, int ldq, const double* qE, const double* vE,
                      int nmodes, const int* order_TTE, const int* order_TEE, const int* order_EEE);
/* Drop the E setup (the reference's do_polarisation = 0). */
int mcmb_lspace_clear_E(mcmb_ctx* ctx);
/* E_lmin, E_lmax, mode count and generation of the E setup (outputs may be NULL). */
int mcmb_lspace_E_info(mcmb_ctx* ctx, int* E_lmin, int* E_lmax, int* nmodes, long* generation);
/* Component pol (MCMB_POL_*) of the Gauss-Legendre Gram, calculate_orthol
 * (orthogonalisation_late.c:239-336, M^E of calculate_orthoMij_E 131-157):
 * with M^X_ij(x_s) = sum_{l in X's range} (2l+1) v_X(l) P_l(x_s) q^X_i q^X_j,
 *   TTT, EEE: sum_s w_s/(48 pi) perm(M_{m_a n_b}), one field;
 *   TTE: sum_s w_s/(16 pi) (M^T_{m0 n0} M^T_{m1 n1} + M^T_{m1 n0} M^T_{m0 n1}) M^E_{m2 n2};
 *   TEE: sum_s w_s/(16 pi) M^T_{m0 n0} (M^E_{m1 n1} M^E_{m2 n2} + M^E_{m2 n1} M^E_{m1 n2}).
 * npt <= 0: 3 max(T lmax, E lmax)/2 + 1 (init_orthol_glint with
 * polarisation); npt must be exact for the component's degree.  n x n with n
 * = the l-space mode count (TTT) or the E setup's (the others).  When the
 * pair products of the bases have a low numerical rank (see
 * mcmb_last_basis_rank) the node sum runs over the compressed bases, the
 * same Gram to rounding; MCMB_GL_PATH=direct keeps the direct node sum. */
int mcmb_gamma_gl_pol(mcmb_ctx* ctx, int pol, int npt, double* gamma_host, double* gamma_dev);

/* The same GL Gram into the caller's device buffer gamma_dev (n x n doubles)
 * without allocating: no hipMalloc, hipFree or allocating library call.  The
 * node table, Legendre table, pair products and M live in one scratch array
 * per GPU (128 MiB, held until the process ends), enough for lmax 2500 with
 * basis width 32 and 3751 nodes; larger setups (more than 4096 nodes or 128
 * MiB of tables) are rejected with MCMB_ERR_ARG.  Every call on a GPU shares
 * that GPU's array, so calls that overlap on one GPU (other threads, other
 * contexts on the same device) race; the call returns when its kernels have
 * finished.  It always takes the direct node sum (mcmb_gamma_gl with
 * MCMB_GL_PATH=direct, bit for bit), since the compressed pair basis of the
 * default path is built in allocated memory; the two paths agree to
 * rounding.  No reference counterpart: for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). */
int mcmb_gamma_gl_noalloc(mcmb_ctx* ctx, int npt, double* gamma_dev);

/* lmin, lmax, mode count and a generation number that changes on every
 * successful mcmb_lspace_set (any output pointer may be NULL). */
int mcmb_lspace_info(mcmb_ctx* ctx, int* lmin, int* lmax, int* nmodes, long* generation);

/* Number of tetrapyd points / rows of the current l-space setup. */
long mcmb_lspace_npoints(mcmb_ctx* ctx);

/* Expand the device point enumeration (rows of T in walk order) into
 * (i,j,k) triples; writes min(cap, N) triples, returns N or a negative status.
 * For bit-exact parity with get_ijk_next_XXX (tetrapyd_tools.c:104-142). */
long mcmb_lspace_points(mcmb_ctx* ctx, int* ijk, long cap);
/* The same from walk index p0: writes min(cap, N - p0) triples, returns N. */
long mcmb_lspace_points_at(mcmb_ctx* ctx, long p0, int* ijk, long cap);

/* Per-point weights w_p (walk order) computed on the device, up to cap. */
long mcmb_lspace_weights(mcmb_ctx* ctx, double* w, long cap);

/* ---- cross-Gamma against the projected primordial basis q-tilde --------- */
/* Load q-tilde as read_qtilde leaves it (fixed_data.c:2052-2061, 2246-2290):
 * qt[p][l][x] for p < np (primordial basis index), multipole l = 0..lmax_q
 * (the table's l index is the multipole), x < nx on the increasing grid
 * xvec[nx] (get_qtilde_xvec); the primordial mode table order_prim[nmodes][3]
 * (find_perm_prim); the TT spectra cl[l], nl[l], bl[l], l = 0..lmax_q
 * (get_cl_TT, get_noise_TT, get_beam_TT); deltaphi (model:delta_phi).  The
 * late side is the l-space setup (mcmb_lspace_set: late basis by multipole,
 * late mode table = find_perm_late_TTT, range [eflag_T_lmin, eflag_T_lmax]). */
int mcmb_qtilde_set(mcmb_ctx* ctx, int np, int lmax_q, int nx, const double* qt, const double* xvec,
                    int nmodes, const int* order_prim, const double* cl, const double* nl,
                    const double* bl, double deltaphi);
/* np, lmax_q, nx, nmodes and generation of the q-tilde setup (outputs may be NULL). */
int mcmb_qtilde_info(mcmb_ctx* ctx, int* np, int* lmax_q, int* nx, int* nmodes, long* generation);
/* Copy the x grid (nx doubles). */
int mcmb_qtilde_xvec(mcmb_ctx* ctx, double* xvec);
/* Gamma_3D of gamma_3d_host (gamma_mpi.c:30-152, calculate_gamma_TTT_3D_nested
 * calculate_gamma_3d.c:274-440, calculate_sijk_TTT 813-860): the rank's share
 * of the walk of T(2, lmax_q) split as decompose_gamma_TTT_3D_mpi splits its
 * point count, z = permsix h^2 lw lw lw with lw(l) = (2l+1)^(-1/3) (h^2 by the
 * l-space setup's h2_mode; the reference's is MCMB_H2_ASYMPTOTIC), late mode
 * products (0 unless all legs are in the late range), and the x integral by
 * the trapezium rule of q3 = (2l+1)^(1/6)/sqrt(C_l + N_l/b_l^2) q-tilde with
 * samples x < 10 dropped where |y| > 1e-18; scaled by 6 dphi^2.  Gamma is
 * n_late x n_prim (rows late modes), written row-major to gamma_host and/or
 * the device buffer gamma_dev; the shares of all ranks sum to rank 0 of 1. */
int mcmb_gamma_cross_3d(mcmb_ctx* ctx, int rank, int nranks, double* gamma_host, double* gamma_dev);
/* Gamma_2D of calculate_gamma_2d (spline MCMB_SPLINE_FMM) or calculate_gamma
 * (MCMB_SPLINE_NATURAL) for every (m, n) at once (calculate_gamma.c:48-303,
 * 511-555, 772-1016): M[p][q][x][s] = sum_l P_l(mu_s) (2l+1)^(5/6) b_l /
 * sqrt(b_l^2 C_l + N_l) qt[p][l][x] b_q(l) over the late range, npt
 * Gauss-Legendre nodes (npt <= 0: 3 lmax/2 + 1), per x the node sum of the
 * 3 x 3 permanents, the same x < 10 rule, the spline integral in x, scaled by
 * dphi^2/(8 pi).  Needs lmax_q >= the late lmax. */
int mcmb_gamma_cross_gl(mcmb_ctx* ctx, int spline, int npt, double* gamma_host, double* gamma_dev);
/* E-mode q-tilde of the polarised cross-Gamma (read_qtilde's E files,
 * fixed_data.c:2052-2063, 2330-2440): qt_E[p][l][x] of the T q-tilde's shape,
 * and per multipole 0..lmax_q the spectra C_TE, C_EE, noise N_TE, N_EE, beams
 * b_TE, b_EE and fsky_E / fsky_T.  Call after mcmb_qtilde_set (whose C_l, N_l,
 * b_l are the TT ones); a new mcmb_qtilde_set drops it. */
int mcmb_qtilde_set_E(mcmb_ctx* ctx, const double* qt_E, const double* cl_TE, const double* cl_EE,
                      const double* nl_TE, const double* nl_EE, const double* bl_TE, const double* bl_EE,
                      double fsky_ratio);
/* fsky ratio and generation of the E-mode q-tilde setup (outputs may be NULL). */
int mcmb_qtilde_E_info(mcmb_ctx* ctx, double* fsky_ratio, long* generation);
/* Component pol (MCMB_POL_*) of the GL cross-Gamma, calculate_gamma_2d /
 * calculate_gamma result[pol] (calculate_gamma.c:48-352, 511-1016): late slot
 * b of the component's mode table reads M^T or M^E by its field (TTT: T T T,
 * TTE: T T E, TEE: T E E, EEE: E E E), a_ab = M^{X_b}[p_a][q_b] and the node
 * sum of the 3 x 3 permanent as above, with
 *   M^E[x][p][q][s] = sum_{l in E range} P_l(mu_s) EEw(l) (s1 qt_E - s2 qt_T)[p][l][x] b^E_q(l),
 *   s1 = b_TT^2 C_TT + N_TT, s2 = sqrt(fsky_E/fsky_T) b_TE^2 C_TE + N_TE,
 *   EEw = (2l+1)^(5/6) b_EE / sqrt(s1 (s1 s4 - s3^2)),
 *   s3 = b_TT b_EE C_TE + N_TE, s4 = b_EE^2 C_EE + N_EE;
 * npt <= 0: 3 max(T lmax, E lmax)/2 + 1.  Gamma is n_late x n_prim with
 * n_late the l-space mode count (TTT) or the E setup's (the others). */
int mcmb_gamma_cross_gl_pol(mcmb_ctx* ctx, int pol, int spline, int npt, double* gamma_host, double* gamma_dev);
/* Per-sample weights c[x] of the x rules above on the grid x[n] (host only):
 * sum_x c[x] y[x] is the rule's integral of y over [x_0, x_{n-1}]. */
int mcmb_spline_weights(int kind, int n, const double* x, double* c);

/* ---- FP64 orthonormalisation on the GPU (ortho_prim_mpi.c:198-284) ------- */
/* norm = sqrt(diag); C = D^-1 Gamma D^-1 = L L^T; ortho = symmetrised D L,
 * lambda = symmetrised (D L)^-1 (update_ortho/update_lambda layouts,
 * fixed_data.c:3285-3295, 3576-3586).  Host arrays, n x n.  Returns
 * MCMB_ERR_NOT_SPD (and the failing pivot in *bad_pivot if non-NULL) when C is
 * not positive definite in FP64. */
int mcmb_orthonormalise(mcmb_ctx* ctx, int n, const double* gamma, double* ortho,
                        double* lambda, double* norm, int* bad_pivot);
/* beta = Gamma^-1 alpha for nrhs right-hand sides (column j = alpha + j*n),
 * with the symmetrised D L factor from mcmb_orthonormalise
 * (gsl_linalg_cholesky_solve, eigenmode_mpi.c:211). */
int mcmb_cholesky_solve(mcmb_ctx* ctx, int n, const double* ortho, int nrhs,
                        const double* alpha, double* beta);

/* Symmetric eigen-decomposition in FP64 on the GPU (parallel cyclic Jacobi),
 * replacing gsl_eigen_symmv + gsl_eigen_symmv_sort(..., GSL_EIGEN_SORT_VAL_DESC)
 * (ortho_prim_mpi.c:223-224, ortho_late_mpi.c:775-776).  a: n x n row-major
 * symmetric.  evals[n] descending.  evecs (n x n, may be NULL): evecs[i*n + k]
 * = component i of eigenvector k (the column layout of GSL's eigenv), with its
 * largest-magnitude component positive.  *sweeps (may be NULL) = sweeps used. */
int mcmb_eigen_sym(mcmb_ctx* ctx, int n, const double* a, double* evals, double* evecs, int* sweeps);

/* Late-side orthonormalisation with the PCA fallback (ortho_late_mpi.c:
 * 752-842): norm = sqrt(diag Gamma), C = D^-1 Gamma D^-1, eigenvalues evals[n]
 * of C (descending).  PCA branch when pca_flag == 0 (the reference's
 * gflag_pca == 0) or the smallest eigenvalue is negative: r = number of
 * positive eigenvalues, lambda[i][j] = V[j][i] / (norm_j sqrt(e_i)) for i < r,
 * lambdainv[i][j] = norm_i V[i][j] sqrt(e_j) for j < r, zero elsewhere (V as
 * evecs above).  Cholesky branch otherwise: lambdainv = D L and lambda =
 * (D L)^-1, lower triangles, zero above.  *rank = r (n in the Cholesky
 * branch).  n x n host arrays. */
int mcmb_orthonormalise_late(mcmb_ctx* ctx, int n, const double* gamma, int pca_flag, double* lambda,
                             double* lambdainv, double* norm, double* evals, int* rank);

/* ---- primordial k-space: Gram and shape projection ------------------------ */
/* Gauss-Legendre k-tetrapyd: N ascending nodes kn[] with weights kw[], basis
 * q[a][0..ldq) at node a, mode order[n][3].  Points are the node-index triples
 * a <= b <= c with kn[c] <= kn[a] + kn[b] (FP64 comparison), multiplicity
 * permsix(a,b,c), weight w_a w_b w_c / (k_a + k_b + k_c): the inner product of
 * calculate_ortho / calculate_eigen (orthogonalisation.c:14-86,
 * calculate_eigen.c:86-145) on the GL grid instead of the uniform grid. */
int mcmb_kspace_set(mcmb_ctx* ctx, int N, const double* kn, const double* kw, int ldq,
                    const double* q, int nmodes, const int* order);
long mcmb_kspace_npoints(mcmb_ctx* ctx);
/* Gram of the k-space modes (calculate_ortho, orthogonalisation.c:14-86, full
 * symmetric n x n instead of the packed triangle); precision FP64 or MIXED. */
int mcmb_kgram(mcmb_ctx* ctx, int precision, double* gamma_host, double* gamma_dev);
/* Projection of the feature shape (calculate_shape.c:933-946, model 4) for
 * ntheta parameter pairs theta[b] = (omega, phi), phi in radians (the
 * reference's a2 percent-of-2pi phase is phi = 2 * 3.1415927 * a2 / 100):
 *   alpha[b][n]      = <S(theta_b), Q_n>      (host, ntheta x n)
 *   dalpha[b][0][n]  = d alpha / d omega, dalpha[b][1][n] = d alpha / d phi
 *                      (tangent mode, same pass; NULL to skip)
 * precision: MCMB_FP64 (FP64 MFMA) or MCMB_FP32 (FP32 MFMA k-sum, FP64 mode
 * sums).  Replaces calculate_eigen's alpha (eigenmode_mpi.c:175); beta follows
 * from mcmb_cholesky_solve with the k-space Gram. */
int mcmb_project_fwd(mcmb_ctx* ctx, int precision, int ntheta, const double* theta,
                     double* alpha, double* dalpha);
/* Adjoint: dtheta[b][0..1] = sum_n gbar[b][n] d alpha_n / d (omega, phi),
 * computed in reverse mode (gbar scattered on the mode triples and pulled back
 * through the mode sums to the k-sum), no Jacobian formed. */
int mcmb_project_bwd(mcmb_ctx* ctx, int precision, int ntheta, const double* theta,
                     const double* gbar, double* dtheta);

/* ---- primordial k-space on the uniform grid (reference parity) ----------- */
/* Uniform k-grid tetrapyd of calculate_ortho / calculate_eigen
 * (orthogonalisation.c:14-86, calculate_eigen.c:86-145): grid points
 * 0..max with k = kmax * index / max (eigenmode_mpi.c:56-59), the walk of
 * get_ijk_next_prim from (min,min,min), point weight
 * calculate_weight(min,max,i,j,k) * dk^3 / (k1+k2+k3) (calculate_volume.c:
 * 12-132; 0 at the origin), basis q[index][0..ldq) (create_basis_prim layout,
 * fixed_data.c:1670-1715), modes order[n][3]. */
int mcmb_kuniform_set(mcmb_ctx* ctx, int min, int max, double kmax, int ldq, const double* q,
                      int nmodes, const int* order);
/* min, max, mode count, kmax, generation of the uniform setup (outputs may be NULL). */
int mcmb_kuniform_info(mcmb_ctx* ctx, int* min, int* max, int* nmodes, double* kmax, long* generation);
long mcmb_kuniform_npoints(mcmb_ctx* ctx);
/* Point weights in walk order (up to cap); returns the point count. */
long mcmb_kuniform_weights(mcmb_ctx* ctx, double* w, long cap);
/* Gram over walk points [p0, p1) (p1 < 0: to the end), full symmetric n x n,
 * FP64 or MIXED (calculate_ortho's sum, orthogonalisation.c:14-86). */
int mcmb_kuniform_gram(mcmb_ctx* ctx, int precision, long p0, long p1, double* gamma_host,
                       double* gamma_dev);
/* alpha[b][n] over walk points [p0, p1) for nshape feature shapes
 * theta[b] = (omega, phi), phi in radians (calculate_eigen with shape3 of
 * model 4, calculate_eigen.c:86-145, calculate_shape.c:933-946); FP64. */
int mcmb_kuniform_project(mcmb_ctx* ctx, int nshape, const double* theta, long p0, long p1,
                          double* alpha);

/* Read-only views of the host tables of the uniform k-grid / l-space setups
 * (valid until the next mcmb_kuniform_set / mcmb_lspace_set; outputs may be
 * NULL).  No GPU work. */
int mcmb_kuniform_tables(mcmb_ctx* ctx, const double** q, int* ldq, const int** order, int* nmodes,
                         int* max);
int mcmb_lspace_tables(mcmb_ctx* ctx, const double** q, int* ldq, const int** order, int* nmodes,
                       int* lmin, int* lmax);
int mcmb_lspace_E_tables(mcmb_ctx* ctx, const double** q, int* ldq, const int** order_TTE,
                         const int** order_TEE, const int** order_EEE, int* nmodes, int* E_lmin, int* E_lmax);

/* ---- Gamma post-processing (gamma_mpi.c:260-295, 801-840) --------------- */
/* orthoinv = lambda^T lambda (orthoinv[i][j] = sum_k lambda[k][i] lambda[k][j])
 * and gamma_out = orthoinv Gamma, FP64 on the GPU; n x n row-major host
 * arrays, orthoinv may be NULL. */
int mcmb_gamma_post(mcmb_ctx* ctx, int n, const double* lambda, const double* gamma, double* orthoinv,
                    double* gamma_out);

/* ---- file layouts (host only) --------------------------------------------- */
/* array_write / array_read (arrays.c:6-21, 54-67) as used by output_gamma,
 * output_ortho, output_lambda, output_eigen (fixed_data.c:3077-3091,
 * 3297-3310, 3588-3601, 2899-2928): raw native-endian FP64, no header.
 * Errors (MCMB_ERR_IO, including a short read) are returned, not exits. */
int mcmb_write_raw(const char* path, long count, const double* data);
int mcmb_read_raw(const char* path, long count, double* data);
/* Mode order table as load_three_int reads it (io.c:229-258): one triple per
 * line, '#' lines skipped; at most cap triples into order[3*cap]. */
int mcmb_load_order(const char* path, int* order, int cap, int* nmodes);

/* ---- file names and q-tilde files (host only) ---------------------------- */
#define MCMB_FILE_EIGEN      0   /* basis:eigen_file   */
#define MCMB_FILE_ORTHO      1   /* basis:ortho_file   */
#define MCMB_FILE_LAMBDA     2   /* basis:lambda_file  */
#define MCMB_FILE_ORTHOL     3   /* basis:orthol_file  (pol: 0 TTT, 1 TTE, 2 TEE, 3 EEE) */
#define MCMB_FILE_LAMBDAL    4   /* basis:lambdal_file (pol as above) */
#define MCMB_FILE_GAMMA      5   /* basis:gamma_file   (pol as above) */
#define MCMB_FILE_PROJ_DATA  6   /* basis:proj_file, data (pol: 0 T, 1 E) */
#define MCMB_FILE_PROJ_SIZE  7   /* basis:proj_file, size file (pol: 0 T, 1 E) */
/* The file name the reference's drivers build from the .ini (parameter.c:
 * 244-257, 317-760): base as given if it starts with '/' (gamma: + "_<pol>",
 * proj: + "_T" / "_E", size files + "_T_size" / "_E_size"; the late-side
 * orthol / lambdal have a name for TTT only), else directory + base + the
 * suffix chain, e.g. gamma TTT: "_TTT_<T_lmin>-<T_lmax>_<order_prim>_
 * <order_late>_<alpha_max>".  MCMB_ERR_ARG if it does not fit in cap. */
int mcmb_file_name(char* out, size_t cap, int kind, int pol, const char* directory, const char* base,
                   int T_lmin, int T_lmax, int E_lmin, int E_lmax, int order_prim, int order_late,
                   int alpha_max);
/* q-tilde files as output_qtilde writes and read_qtilde reads them
 * (fixed_data.c:2076-2330): size_file holds {lsize + 1, xsize + 1} as native
 * ints; per primordial index p < np the file data_file + "_<p>" (order4: the
 * last two are "_l1", "_l2") holds one (lsize+1) x (xsize+1) FP64 array with
 * the x grid in row 0, the multipoles lvec in column 0 and qt[p][l][x].
 * mcmb_read_qtilde returns the sizes alone when lvec / xvec / qt are NULL,
 * and MCMB_ERR_ARG when they exceed cap_l / cap_x; qt is [np][lsize][xsize]. */
int mcmb_write_qtilde(const char* size_file, const char* data_file, int np, int order4, int lsize, int xsize,
                      const int* lvec, const double* xvec, const double* qt);
int mcmb_read_qtilde(const char* size_file, const char* data_file, int np, int order4, int cap_l, int cap_x,
                     int* lsize, int* xsize, int* lvec, double* xvec, double* qt);
/* create_basis_late's T basis of eflag_order_late 9 (fixed_data.c:1923-1950):
 * basis[i][0..pmax-2] = P_j((Tvec[i] - Tmin)/(Tmax - Tmin)) (shifted Legendre,
 * basis_functions_legendre), basis[i][pmax-1] = w/z, basis[i][pmax] = w z
 * with x = l(l+1)/(Tmax(Tmax+1)), y = (2l+1)/(2Tmax+1), z = sqrt(x), w =
 * y^(1/6), l = Tvec[i]; basis is [Tsize][pmax+1] (row i is Tvec[i], the
 * reference's row l - Tmin). */
int mcmb_basis_late_order9(int Tsize, int pmax, double Tmin, double Tmax, const double* Tvec, double* basis);

/* ---- quadrature ----------------------------------------------------------- */
/* n-node Gauss-Legendre rule on [-1,1], nodes ascending (host only, no GPU). */
int mcmb_gauss_legendre(int n, double* x, double* w);

/* ---- timing of the last Gram launch sequence ------------------------------ */
/* Milliseconds spent in the dominant kernel of the last mcmb_gamma_tet /
 * mcmb_gamma_gl call (HIP events on the context stream) and its launch count. */
int mcmb_last_kernel_ms(mcmb_ctx* ctx, double* ms, int* launches);
/* ---- device buffers for callers without their own allocator ------------- */
/* Allocate / free device memory on the context's GPU, and copy a device buffer
 * to the host (ordered after the context stream's work).  No reference
 * counterpart: they let a C caller keep the Gram in HBM (mcmb_gamma_tet with a
 * device output) without linking HIP itself. */
int mcmb_device_alloc(mcmb_ctx* ctx, size_t bytes, void** dev);
int mcmb_device_free(mcmb_ctx* ctx, void* dev);
int mcmb_device_to_host(mcmb_ctx* ctx, void* host, const void* dev, size_t bytes);
/* Copy bytes (a multiple of 16, 16-byte aligned buffers) from src to dst on
 * the device with a plain copy kernel on the context stream; no reference
 * counterpart (a known-size read + write to calibrate HBM byte counters). */
int mcmb_device_copy(mcmb_ctx* ctx, void* dst, const void* src, size_t bytes);

/* Per-phase milliseconds of the last call that records phases (the FP64
 * separable Gram: basis compression on the first call of a setup, pair
 * products or weights, l3 / l2 / l1 contractions, expansion, assembly),
 * each between HIP events on the context stream.  Writes up to cap entries of
 * ms[] and names[] (static strings) and returns the phase count (0 if the last
 * such call recorded none; with MCMB_PHASES=0 in the environment no call
 * records them, which saves about 25 us per Gram at 4960 modes).  No
 * reference counterpart. */
int mcmb_last_phases(mcmb_ctx* ctx, int cap, double* ms, const char** names);
/* Host wall time (ms) of the parts of the last separable Gram call, in call
 * order: basis factorisation, the row tables' parts (row runs and blocks, row
 * groups and tables on the host, table allocation, table upload), buffers
 * (each only when the call built it), launch to completion and the whole
 * call; returns the count. */
int mcmb_last_setup(mcmb_ctx* ctx, int cap, double* ms, const char** names);
/* *rank = the basis columns the last FP64 separable or Gauss-Legendre Gram
 * contracted over: r when the pair products q_x q_y had numerical rank r <=
 * NP / 2 (NP = R(R+1)/2) and were replaced by an orthonormal basis of their
 * span, 0 when the NP pair products themselves were contracted (for the GL
 * Gram: the direct node sum), -1 when the last Gram took another contraction
 * order.  For the polarised GL components (mcmb_gamma_gl_pol) it is the rank
 * of the basis on two legs (T for TTE, E for TEE and EEE); the mixed
 * components contract over both bases only when both have such a rank.  No
 * reference counterpart. */
int mcmb_last_basis_rank(mcmb_ctx* ctx, int* rank);

/* ---- reference-named entry points (legacy shims over the bound context) --- */
/* Bind the context the legacy symbols use (not owned). */
int mcmb_legacy_bind(mcmb_ctx* ctx);
/* Status of the last void legacy call (gamma_3d_host, get_next_ijk_bound). */
int mcmb_legacy_status(void);

/* tetrapyd_tools.h:53-65,98-111 */
int  get_ijk_next_XXX(int lmax, int* i, int* j, int* k);
void check_ijk_min_XXX(int lmin, int* i, int* j, int* k);
void check_ijk_max_XXX(int lmax, int* i, int* j, int* k);
void decompose_tetrapyd_XXX(tetrapyd_limits* thread_limits, int rank, int numranks,
                            int nthreads, int lmin, int lmax);
int  get_ijk_next_prim(int max, int* i, int* j, int* k);
void check_ijk_min_prim(int min, int* i, int* j, int* k);
void check_ijk_max_prim(int max, int* i, int* j, int* k);
void decompose_tetrapyd_prim(tetrapyd_limits* thread_limits, int rank, int numranks,
                             int nthreads, int min, int max);
void divide_tasks(long int tasks, int nranks, long int* tasks_per_rank);
/* tetrapyd_tools.h:81-84 (tetrapyd_tools.c:226-414): the mixed walk, i <= j
 * over [lminX, lmaxX], k over [lminY, lmaxY], bit-exact with the reference
 * including its quirks (a new i restarts at k = lminY without a parity bump;
 * a j with an empty k range still yields one point; an odd lmaxY ends the
 * walk at (lmaxX - 1, lmaxX, lmaxY)), so it is the mixed tetrapyd exactly
 * when lminX = lminY and lmaxX = lmaxY, both even. */
void decompose_tetrapyd_XXY(tetrapyd_limits* thread_limits, int rank, int numranks, int nthreads, int lminX,
                            int lminY, int lmaxX, int lmaxY);
int  get_ijk_next_XXY(int lminX, int lminY, int lmaxX, int lmaxY, int* i, int* j, int* k);
void check_ijk_max_XXY(int lmaxX, int lmaxY, int* i, int* j, int* k);
void check_ijk_min_XXY(int lminX, int lminY, int* i, int* j, int* k);
/* get_next_ijk_bound with an explicit lsize = lmax + 1 (host only) */
void mcmb_next_ijk_bound(int lsize, int i0, int j0, int k0, int* i1, int* j1, int* k1,
                         int final_i, int final_j, int final_k, int* blocksize,
                         int maxblocksize);
/* calculate_gamma_3d.c:92-211 with explicit lsize and thread count (host
 * only): fills td[0 .. nthreads) with each thread's walk bounds. */
void mcmb_decompose_gamma_TTT_3D(int lsize, int rank, int numranks, double workshare, int nthreads,
                                 thread_g3d_data_t* td);
/* calculate_gamma_3d.h:46 -- the same over the bound context (lsize =
 * get_qtilde_T_lsize() of its q-tilde setup, else lmax + 1 of its l-space
 * setup), nthreads = omp_get_num_threads() when the program links OpenMP,
 * else 1. */
void decompose_gamma_TTT_3D_mpi(thread_g3d_data_t* thread_data, int rank, int numranks, double workshare);
/* calculate_gamma_3d.h:72-73 (lsize as decompose_gamma_TTT_3D_mpi takes it) */
void get_next_ijk_bound(int i0, int j0, int k0, int* i1, int* j1, int* k1,
                        int final_i, int final_j, int final_k, int* blocksize,
                        int maxblocksize);
/* fixed_data.c:4244 */
double permsix(int i, int j, int k);
/* calculate_geometric.c:59 */
double calculate_geometric(int l1, int l2, int l3);
/* basis_functions.c:88 */
int basis_functions_legendre(double** array, int size, int pmax, double min, double max,
                             double* values);
/* gamma_mpi.c:30 -- fills the caller's row-major gamma_size x gamma_size
 * buffer (zeroed first) with this rank's share of the 3D cross-Gamma of the
 * bound context (mcmb_gamma_cross_3d: late basis of the l-space setup against
 * the q-tilde setup, 6 dphi^2 included; numranks <= 1: the whole walk);
 * summing the buffers of all ranks gives the whole, as the MPI_Reduce at
 * gamma_mpi.c:764 does.  workshare (the Xeon Phi offload split) is accepted
 * and ignored.  Without a q-tilde setup the buffer stays zero and
 * mcmb_legacy_status() is MCMB_ERR_STATE. */
void gamma_3d_host(double* gamma_flat, int gamma_size, int rank, int numranks,
                   double workshare);
/* calculate_gamma.c:773 / 936 -- result[0] = Gamma_TTT(m, n) of the GL
 * cross-Gamma (FMM spline in x / natural spline in x) of the bound context,
 * computed whole on the first call and cached until a setup changes; safe to
 * call from several threads at once (the first call fills the cache under a
 * lock).  result[0] alone, or result[0..3] = TTT, TTE, TEE, EEE of
 * mcmb_gamma_cross_gl_pol when mcmb_lspace_set_E and mcmb_qtilde_set_E are
 * both bound (the reference's do_polarisation = 1). */
void calculate_gamma_2d(int m, int n, double* result);
void calculate_gamma(int m, int n, double* result);
/* calculate_gamma.c:48, 255, 354, 425 -- the M_ij tables are formed on the
 * GPU inside the first calculate_gamma(_2d); these check the setups
 * (precompute_gammaMij_T also forms the FMM-spline Gamma), sync_gammaMij has
 * nothing to sum on one GPU. */
void init_gamma_glint(void);
void precompute_gammaMij_T(void);
void calculate_gammaMij_T(int i, int j);
void sync_gammaMij(void);
/* calculate_gamma.c:306, 388 -- the E tables come with the polarised
 * cross-Gamma (mcmb_lspace_set_E and mcmb_qtilde_set_E bound; MCMB_ERR_STATE
 * in mcmb_legacy_status otherwise) */
void precompute_gammaMij_E(void);
void calculate_gammaMij_E(int i, int j);
/* fixed_data.c get_qtilde_xsize / get_qtilde_xvec / get_qtilde_T_lsize of the
 * bound q-tilde setup (lsize = lmax_q + 1) */
int get_qtilde_xsize(void);
int get_qtilde_xvec(double* xvec);   /* returns 0, or the status without a setup */
int get_qtilde_T_lsize(void);
/* orthogonalisation_late.c:23, 99, 159 -- as above for the auto-Gram:
 * init_orthol_glint returns 0, calculate_orthoMij_T checks state (every M_ij
 * is formed by the first calculate_orthol), sync_orthoMij has nothing to sum. */
double init_orthol_glint(void);
void calculate_orthoMij_T(int i, int j);
/* orthogonalisation_late.c:131 -- the same for the E field (MCMB_ERR_STATE in
 * mcmb_legacy_status without mcmb_lspace_set_E) */
void calculate_orthoMij_E(int i, int j);
void sync_orthoMij(void);
/* orthogonalisation_late.c:338 -- result[0] = Gamma_mn (TTT) of the bound
 * context's cached FP64 tetrapyd Gram (computed on first use). */
int calculate_orthol_3D(int m, int n, double* result);
/* orthogonalisation_late.c:239 -- result[0] = Gamma_mn from the cached
 * Gauss-Legendre-factorised Gram; with an E setup bound (the reference's
 * do_polarisation = 1), result[0..3] = TTT, TTE, TEE, EEE of
 * mcmb_gamma_gl_pol, all on 3 max(T lmax, E lmax)/2 + 1 nodes. */
int calculate_orthol(int m, int n, double* result);
/* orthogonalisation.c:14 -- adds the FP64 Gram of the share of the uniform
 * walk covered by block[0 .. nthreads) of the last decompose_tetrapyd_prim
 * call into the packed upper triangle ortho[n(n+1)/2] (r <= s, row major),
 * using the bound context's mcmb_kuniform_set tables (same min, max).  Every
 * block is integrated; the reference's threaded build integrates block 0
 * only (its OpenMP pragma is commented out, orthogonalisation.c:34,37). */
void calculate_ortho(int min, int max, tetrapyd_limits* block, double* ortho);
/* calculate_eigen.c:86 -- adds alpha[n] of the feature shape (model 4:
 * params.a1 = omega, params.a2 = phase in percent of 2 pi) over the same
 * share of the uniform walk. */
void calculate_eigen(int min, int max, shape_params params, tetrapyd_limits* block, double* alpha);
/* orthogonalisation.c:279 -- pijk of the bound uniform k-grid setup */
double pijk(int r, int i, int j, int k);
/* orthogonalisation_late.c:620 -- plijk_TTT of the bound l-space setup
 * (basis row = multipole; 0 outside [lmin, lmax]) */
double plijk_TTT(int r, int l1, int l2, int l3);
/* orthogonalisation_late.c:646-715 -- the mixed mode products of the bound
 * T and E setups (legs T, T, E / T, E, E / E, E, E; 0 when a leg is outside
 * its field's range) */
double plijk_TTE(int r, int l1, int l2, int l3);
double plijk_TEE(int r, int l1, int l2, int l3);
double plijk_EEE(int r, int l1, int l2, int l3);
/* fixed_data.c:263 */
void find_perm_prim(int r, int* p1, int* p2, int* p3);
/* fixed_data.c:983-990 -- mode r of the bound l-space setup's order table
 * (-1s outside it) */
void find_perm_late_TTT(int r, int* p1, int* p2, int* p3);
/* fixed_data.c:992-1019 -- mode r of the bound E setup's tables, in the slot
 * order given to mcmb_lspace_set_E (-1s outside them) */
void find_perm_late_TTE(int r, int* p1, int* p2, int* p3);
void find_perm_late_TEE(int r, int* p1, int* p2, int* p3);
void find_perm_late_EEE(int r, int* p1, int* p2, int* p3);
/* global.h:429, 730 -- of the bound uniform k-grid setup (0 if unset) */
int get_terms_prim(void);
double get_kmax(void);
/* calculate_shape.c:26, 197 -- the feature shape (model 4) only */
double shape(double k1, double k2, double k3, shape_params params);
double shape3(double k1, double k2, double k3, shape_params params);
/* fixed_data.c:1670 -- Legendre orders: rebuilds the bound uniform k-grid
 * setup on size points of [0, max] with basis P_p(2 vec[i]/max - 1), keeping
 * its order table; returns 0 or a negative status. */
int create_basis_prim(int size, double max, double* vec);
/* io.c:229 -- 1 on success, 0 if the file cannot be read */
int load_three_int(char* filename, int* values, int* size);

#ifdef __cplusplus
}
#endif
#endif
