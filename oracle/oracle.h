/*
 * oracle.h -- CPU FP64 restatement of bfam's DG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker.
 *
 * What it restates (reference file:line, all in src/bfam.c unless noted):
 *   dense helpers   bfam_util_mtranspose/mTmmult/mmmult/lu_factor/lu_solve/
 *                   backslash/forwardslash                     1091-1306
 *   Jacobi/LGL      bfam_jacobi_h_inv_sqrt 2700, _p 2754, grad 2843,
 *                   gauss_quadrature_half 2871, gauss 2996, GLL 3013,
 *                   vandermonde 3042/3052, interpolation 3063,
 *                   differentiation 3082, mass 3101
 *   mortar ops      fill_grid_data 5459, fill_interp_proj_data 5471,
 *                   fill_hanging_data 5516, multiply_projections 5543,
 *                   create_interpolators 5606
 *   maps            gmask_set 6599 (faces), buildmaps 6783,
 *                   orientation tables 3630-3646, glue mapOp 7430-7463
 *   LSRK tables     bfamo_lsrk_init 8996-9090
 *   pcg32           bfam_pcg32_* 8328-8387 (O'Neill's published PCG32)
 * The right-hand side itself is absent from the reference (apps supplied it);
 * here it is specified from published methods:
 *   - strong-form nodal DG on tensor LGL points (Hesthaven & Warburton 2008)
 *   - exact-Riemann (upwind) flux for isotropic velocity-stress elasticity,
 *     split into P (normal) and S (tangential) parts (Wilcox, Stadler,
 *     Burstedde & Ghattas, JCP 229 (2010) 9373, sec. 3); acoustics is the
 *     mu = 0 special case with p = -trace(sigma)/dim
 *   - LSRK 2N-storage update dq = A dq + dt L(q, t + C dt); q += B dq
 *     (Carpenter & Kennedy, NASA TM-109112, 1994)
 */
#ifndef BFAMX_ORACLE_H
#define BFAMX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- dense helpers and operators (all FP64, column-major) ---- */
void orc_jacobi_p(double alpha, double beta, int N, int nx, const double *x,
                  double *P);
void orc_grad_jacobi_p(double alpha, double beta, int N, int nx,
                       const double *x, double *dP);
int orc_jacobi_gauss(double alpha, double beta, int N, double *x, double *w);
int orc_jacobi_gauss_lobatto(double alpha, double beta, int N, double *x,
                             double *w);
void orc_vandermonde(int N, int nx, const double *x, double *V);
void orc_grad_vandermonde(int N, int nx, const double *x, double *V);
/* I = V_x V^{-1}, nx x (N+1) */
void orc_interpolation(int N, int nx, const double *x, const double *V,
                       double *I);
/* D[i + j*(N+1)] = l_j'(r_i) */
void orc_differentiation(int N, const double *r, const double *V, double *D);
void orc_mass(int N, const double *V, double *M);
/* backslash: C = A \ B  (A m x m, B m x n) */
void orc_backslash(int m, int n, const double *A, const double *B, double *C);

/* GLL nodes r, weights w, wi = 1/w, Dr, exactly as generic_init 7087-7155 */
int orc_dgx_ops(int N, double *r, double *w, double *wi, double *Dr);

/*
 * create_interpolators (5606): 1-D operators from order Na to Nb.
 * prj, mass_prj, wi_mass_prj each hold 5 blocks of (Nb+1)*(Na+1) doubles
 * (column-major, rows = Nb side).  prj block 0 is absent (identity, NULL in
 * the reference) when Na == Nb; *prj0_null is set to 1 then and block 0 is
 * left zero.
 */
int orc_interpolator(int Na, int Nb, double *prj, double *mass_prj,
                     double *wi_mass_prj, int *prj0_null);

/* ---- LSRK tables (bfamo_lsrk_init 8996-9090) ---- */
enum
{
  ORC_LSRK_KC54 = 0,
  ORC_LSRK_FE = 1,
  ORC_LSRK_HEUN = 2,
  ORC_LSRK_W33 = 3
};
/* returns the number of stages; A,B have >= 5 entries, C >= 6 */
int orc_lsrk_table(int method, double *A, double *B, double *C);

/* ---- integer maps ---- */
/* faces part of gmask_set (6599-6781): gmask[f*Nfp + n] */
void orc_gmask_faces(int dim, int N, int32_t *gmask);
/* buildmaps (6783-6877).  EToF = orientation*Nfaces + face. */
void orc_buildmaps(int dim, int N, int32_t K, const int32_t *EToE,
                   const int8_t *EToF, int32_t *vmapM, int32_t *vmapP);
/* glue orientation permutations (7430-7463); dim_glue 1 or 2 */
void orc_glue_mapOp(int dim_glue, int N, int32_t *mapOp);
/* p8est orientation code conversion (3630-3656): BFAM_P8EST_ORIENTATION */
int orc_p8est_orientation(int f, int nf, int o);
int orc_p8est_perm_to_order(int bo, int h);
int orc_p8est_perm_to_order_inv(int bo, int h);

/* ---- pcg32 (8328-8387) ---- */
typedef struct
{
  uint64_t state;
  uint64_t inc;
} orc_pcg32_t;
void orc_pcg32_seed(orc_pcg32_t *rng, uint64_t initstate, uint64_t initseq);
uint32_t orc_pcg32_next(orc_pcg32_t *rng);
uint32_t orc_pcg32_bounded(orc_pcg32_t *rng, uint32_t bound);

/* ---- periodic / non-periodic brick in lexicographic element order ---- */
/* element e = ix + nx*(iy + ny*iz); EToF = face of the neighbour (o = 0) */
void orc_brick(int dim, const int *n, const int *periodic, int32_t *EToE,
               int8_t *EToF);

/* ---- right-hand side and time stepping ---- */
enum
{
  ORC_ACOUSTIC2D = 0,
  ORC_ELASTIC3D = 1
};

typedef struct
{
  int physics;
  int N;
  int32_t K;
  const double *r;  /* N+1 */
  const double *w;  /* N+1 */
  const double *Dr; /* (N+1)^2 */
  const int32_t *vmapM;
  const int32_t *vmapP;
  const double *h;   /* K*dim element edge lengths */
  const double *x0;  /* K*dim element lower corners (forcing only) */
  const double *rho; /* K */
  const double *lam; /* K (acoustic: bulk modulus rho c^2) */
  const double *mu;  /* K (acoustic: ignored) */
  /* point source: rate[f][src_elem*Np+n] += src_wt[n]*src_amp[f]*ricker(t) */
  int32_t src_elem; /* -1: none */
  const double *src_wt;
  double src_amp[9];
  double src_f0, src_t0;
  /* manufactured forcing: rate[f] += mms_F[f]*P_f(x)*tau_f(t) */
  int mms;
  double mms_k, mms_omega;
  double mms_F[9];
  /* K*6 flags (elastic): faces whose flux comes from mortars; NULL: none */
  const int8_t *glued;
  /* K*2dim physical boundary kinds of self-mapped faces: 0 none, 1 free
   * surface, 2 rigid wall, 3 absorbing (plus state = mirrored own trace);
   * NULL: none */
  const int8_t *bc;
} orc_problem_t;

int orc_nfields(int physics);
/* q, rhs: nfields blocks of K*Np (bfam field layout, one array per field) */
void orc_rhs(const orc_problem_t *p, double t, const double *q, double *rhs);
/* nsteps of the given LSRK method; dq and work are scratch of the same size */
void orc_lsrk_run(const orc_problem_t *p, int method, double t0, double dt,
                  int nsteps, double *q, double *dq, double *work);
/* ---- several elastic subdomains glued by mortars ---- */
/* one interface piece: face s_face of (s_sub, s_elem) coincides with the
 * mortar; face b_face of (b_sub, b_elem) is the same face (quad = -1) or the
 * big face of a 2:1 hanging interface of which the mortar is quadrant quad
 * (bit 0: upper half along the first tangential axis, bit 1: second) */
typedef struct
{
  int32_t s_sub, s_elem, b_sub, b_elem;
  int8_t s_face, b_face, quad, order; /* order: glue grid N_g, 0 = max */
} orc_mortar_t;
/* q[s], rhs[s]: the field blocks of subdomain s */
void orc_rhs_multi(const orc_problem_t *const *subs, int nsub,
                   const orc_mortar_t *m, int32_t nm, double t,
                   const double *const *q, double *const *rhs);
void orc_lsrk_run_multi(const orc_problem_t *const *subs, int nsub,
                        const orc_mortar_t *m, int32_t nm, int method,
                        double t0, double dt, int nsteps, double *const *q,
                        double *const *dq, double *const *work);
/* 0.5 * sum J w (rho|v|^2 + sigma : C^{-1} : sigma), FP64 with Kahan sums */
double orc_energy(const orc_problem_t *p, const double *q);
/* manufactured solution at time t, and the forcing coefficients */
void orc_mms_exact(const orc_problem_t *p, const double *A, double t,
                   double *q);
void orc_mms_forcing_coef(const double *A, double k, double omega, double rho,
                          double lam, double mu, double *F);
double orc_ricker(double f0, double t0, double t);

int orc_num_threads(void);
void orc_set_num_threads(int n);

/* field transfer after h-adaptation through the transfer maps and the 1-D
 * projection operators (bfam.h:1122-1135, bfam.c:5205-5387, 8399-8404) */
int orc_transfer(int nf, int32_t num_dst, const uint8_t *flags,
                 const int8_t *dst_chld, const int32_t *src_subd,
                 const int32_t *src_elem, const int8_t *c_chld,
                 const int32_t *c_subd, const int32_t *c_elem,
                 const int *src_N, const int32_t *src_K,
                 const double *const *src_q, const int32_t *dst_subd,
                 const int32_t *dst_elem, const int *dst_N,
                 const int32_t *dst_K, double *const *dst_q);

#ifdef __cplusplus
}
#endif
#endif
