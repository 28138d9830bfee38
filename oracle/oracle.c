/*
 * oracle.c -- CPU FP64 restatement of bfam's DG hot path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  It is never linked into or
 * called by the shipped library.  Each function cites the reference lines it
 * follows (src/bfam.c unless stated).  Compiled with -ffp-contract=off so that
 * the operator builders round like the reference's gnu99/-O2 build.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAX(a, b) ((a) > (b) ? (a) : (b))

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* threads of the following parallel loops (n >= 1) */
void orc_set_num_threads(int n)
{
#ifdef _OPENMP
  if (n >= 1)
    omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* BFAM_LONG_REAL_APPROX_EQ (bfam.h:182, bfam.c:23) with long real = double */
static int approx_eq(double x, double y, double K)
{
  const double eps = DBL_EPSILON;
  return fabs(x - y) < eps * eps + K * eps * ORC_MAX(fabs(x), fabs(y));
}

/* ------------------------------------------------------------------------ */
/* dense helpers (1091-1306)                                                */
/* ------------------------------------------------------------------------ */

/* B (n x m, ldb) = A^T where A is m x n with leading dim lda (1091) */
static void mtranspose(int m, int n, const double *A, int lda, double *B,
                       int ldb)
{
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j)
      B[i * ldb + j] = A[j * lda + i];
}

/* C += A^T B, A is k x m (1100) */
static void mTmmult(int m, int n, int k, const double *A, int lda,
                    const double *B, int ldb, double *C, int ldc)
{
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j)
      for (int p = 0; p < k; ++p)
        C[j * ldc + i] += A[i * lda + p] * B[j * ldb + p];
}

/* C += A B (1111) */
static void mmmult(int m, int n, int k, const double *A, int lda,
                   const double *B, int ldb, double *C, int ldc)
{
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j)
      for (int p = 0; p < k; ++p)
        C[j * ldc + i] += A[p * lda + i] * B[j * ldb + p];
}

/* Gaussian elimination with complete pivoting, Golub & van Loan 3.4.2 (1122) */
static void lu_factor(int n, double *A, int *p, int *q)
{
  for (int k = 0; k < n; ++k)
    p[k] = q[k] = k;

  for (int k = 0; k < n - 1; ++k)
  {
    int mu = k, lambda = k;
    double a_max = 0.0;
    for (int j = k; j < n; ++j)
      for (int i = k; i < n; ++i)
      {
        const double a_abs = fabs(A[i + n * j]);
        if (a_abs > a_max)
        {
          a_max = a_abs;
          mu = i;
          lambda = j;
        }
      }

    int t = p[k];
    p[k] = p[mu];
    p[mu] = t;
    for (int j = 0; j < n; ++j)
    {
      const double r = A[k + n * j];
      A[k + n * j] = A[mu + n * j];
      A[mu + n * j] = r;
    }

    t = q[k];
    q[k] = q[lambda];
    q[lambda] = t;
    for (int i = 0; i < n; ++i)
    {
      const double r = A[i + n * k];
      A[i + n * k] = A[i + n * lambda];
      A[i + n * lambda] = r;
    }

    if (fabs(A[k + n * k]) > DBL_EPSILON * DBL_EPSILON)
    {
      for (int i = k + 1; i < n; ++i)
        A[i + n * k] = A[i + n * k] / A[k + n * k];
      for (int i = k + 1; i < n; ++i)
        for (int j = k + 1; j < n; ++j)
          A[i + n * j] = A[i + n * j] - A[i + n * k] * A[k + n * j];
    }
  }
}

/* forward (unit L) then backward substitution, column versions (1202) */
static void lu_solve(int n, const double *LU, double *x, const int *p,
                     const int *q, double *work)
{
  for (int k = 0; k < n; ++k)
    work[k] = x[p[k]];
  for (int j = 0; j < n - 1; ++j)
    for (int i = j + 1; i < n; ++i)
      work[i] = work[i] - work[j] * LU[i + n * j];
  for (int j = n - 1; j > 0; --j)
  {
    work[j] = work[j] / LU[j + n * j];
    for (int i = 0; i < j; ++i)
      work[i] = work[i] - work[j] * LU[i + n * j];
  }
  work[0] = work[0] / LU[0];
  for (int k = 0; k < n; ++k)
    x[q[k]] = work[k];
}

/* C = A \ B, A is m x m, B is m x n (1262) */
void orc_backslash(int m, int n, const double *A, const double *B, double *C)
{
  double *LU = malloc(sizeof(double) * m * m);
  double *work = malloc(sizeof(double) * m);
  int *p = malloc(sizeof(int) * m);
  int *q = malloc(sizeof(int) * m);
  memcpy(LU, A, sizeof(double) * m * m);
  memcpy(C, B, sizeof(double) * m * n);
  lu_factor(m, LU, p, q);
  for (int j = 0; j < n; ++j)
    lu_solve(m, LU, C + j * m, p, q, work);
  free(q);
  free(p);
  free(work);
  free(LU);
}

/* C = A / B = A B^{-1}, A is m x n, B is n x n (1287) */
static void forwardslash(int m, int n, const double *A, const double *B,
                         double *C)
{
  double *AT = malloc(sizeof(double) * n * m);
  double *BT = malloc(sizeof(double) * n * n);
  double *CT = malloc(sizeof(double) * n * m);
  mtranspose(m, n, A, m, AT, n);
  mtranspose(n, n, B, n, BT, n);
  orc_backslash(n, m, BT, AT, CT);
  mtranspose(n, m, CT, n, C, m);
  free(CT);
  free(BT);
  free(AT);
}

/* ------------------------------------------------------------------------ */
/* Jacobi polynomials and quadrature (2681-3135)                            */
/* ------------------------------------------------------------------------ */

/* {h_N^{(a,b)}}^{-1/2}, Szego (4.3.3) (2700) */
static double jacobi_h_inv_sqrt(double alpha, double beta, int N)
{
  double lgn = -(alpha + beta + 1) * log(2) - lgamma(N + alpha + 1) -
               lgamma(N + beta + 1) + log(2 * N + alpha + beta + 1) +
               lgamma(N + 1) + lgamma(N + alpha + beta + 1);
  return sqrt(exp(lgn));
}

/* orthonormal Jacobi polynomial by the three-term recurrence (2754) */
void orc_jacobi_p(double alpha, double beta, int N, int nx, const double *x,
                  double *P)
{
  for (int i = 0; i < nx; ++i)
  {
    double P_n_2;
    double P_n_1 = 1;
    double P_n_0 = ((alpha + beta + 2) / 2) * x[i] + (alpha - beta) / 2;
    if (N == 0)
      P[i] = P_n_1;
    else if (N == 1)
      P[i] = P_n_0;
    else
    {
      for (int n = 2; n < N + 1; ++n)
      {
        double a = (2 * n + alpha + beta - 1) * (2 * n + alpha + beta) /
                   (2 * n * (n + alpha + beta));
        double b = (beta * beta - alpha * alpha) * (2 * n + alpha + beta - 1) /
                   (2 * n * (n + alpha + beta) * (2 * n + alpha + beta - 2));
        double c = (n + alpha - 1) * (n + beta - 1) * (2 * n + alpha + beta) /
                   (n * (n + alpha + beta) * (2 * n + alpha + beta - 2));
        P_n_2 = P_n_1;
        P_n_1 = P_n_0;
        P_n_0 = (a * x[i] - b) * P_n_1 - c * P_n_2;
      }
      P[i] = P_n_0;
    }
  }
  const double h = jacobi_h_inv_sqrt(alpha, beta, N);
  for (int i = 0; i < nx; ++i)
    P[i] *= h;
}

/* d/dx p_N = sqrt(N(N+a+b+1)) p_{N-1}^{(a+1,b+1)} (2843) */
void orc_grad_jacobi_p(double alpha, double beta, int N, int nx,
                       const double *x, double *dP)
{
  if (N == 0)
  {
    for (int i = 0; i < nx; ++i)
      dP[i] = 0.0;
    return;
  }
  orc_jacobi_p(alpha + 1, beta + 1, N - 1, nx, x, dP);
  const double scale = sqrt(N * (N + alpha + beta + 1));
  for (int i = 0; i < nx; ++i)
    dP[i] *= scale;
}

/*
 * One half of the Gauss nodes of P_{N+1}: Gatteschi-Pittaluga initial guess
 * (Hale & Townsend, SISC 35 (2013) eq. 3.19) then Newton in theta (2871).
 * Returns 0 on success, -1 if Newton does not converge in 200 iterations
 * (the reference aborts there).
 */
static int gauss_half(double alpha, double beta, int N, int half, double *x,
                      double *w)
{
  const int MAX_ITERATIONS = 200;
  const int nk = half ? (((N + 1) % 2) ? (N + 1) / 2 + 1 : (N + 1) / 2)
                      : (N + 1) / 2;
  if (nk == 0)
    return 0;

  const double tworho = 2 * (N + 1) + alpha + beta + 1;
  const double pi = 4 * atan(1.0);
  double *th0 = malloc(sizeof(double) * nk);
  double *th1 = malloc(sizeof(double) * nk);
  double *p0 = malloc(sizeof(double) * nk);
  double *dp0 = malloc(sizeof(double) * nk);

  for (int k = nk; k > 0; --k)
  {
    const int khat = half ? nk - k : k - 1;
    const double phik = (2 * k + alpha - 0.5) * pi / tworho;
    th1[khat] = phik + 1 / (tworho * tworho) *
                           ((0.25 - alpha * alpha) * 1 / tan(0.5 * phik) -
                            (0.25 - beta * beta) * tan(0.5 * phik));
  }

  int converged = 0;
  for (int it = 0; it < MAX_ITERATIONS; ++it)
  {
    double *t = th0;
    th0 = th1;
    th1 = t;
    for (int k = 0; k < nk; ++k)
      x[k] = cos(th0[k]);
    orc_jacobi_p(alpha, beta, N + 1, nk, x, p0);
    orc_grad_jacobi_p(alpha, beta, N + 1, nk, x, dp0);
    for (int k = 0; k < nk; ++k)
      th1[k] = th0[k] - p0[k] / (-sin(th0[k]) * dp0[k]);
    int diff = 0;
    for (int k = 0; k < nk; ++k)
      diff += !approx_eq(th0[k], th1[k], 10);
    if (!diff)
    {
      converged = 1;
      break;
    }
  }

  if (converged)
  {
    for (int k = 0; k < nk; ++k)
      x[k] = cos(th1[k]);
    orc_grad_jacobi_p(alpha, beta, N + 1, nk, x, dp0);
    for (int k = 0; k < nk; ++k)
    {
      const double s = sin(th1[k]);
      w[k] = tworho / (s * s * dp0[k] * dp0[k]);
    }
  }
  free(th0);
  free(th1);
  free(p0);
  free(dp0);
  return converged ? 0 : -1;
}

/* Gauss nodes/weights of P_{N+1}^{(a,b)} (2996) */
int orc_jacobi_gauss(double alpha, double beta, int N, double *x, double *w)
{
  const int nk_floor = (N + 1) / 2;
  if (gauss_half(alpha, beta, N, 1, x + nk_floor, w + nk_floor))
    return -1;
  if (gauss_half(beta, alpha, N, 0, x, w))
    return -1;
  for (int k = 0; k < nk_floor; ++k)
    x[k] *= -1;
  return 0;
}

/* Gauss-Lobatto: interior = Gauss of (a+1,b+1), weights from P_N (3013) */
int orc_jacobi_gauss_lobatto(double alpha, double beta, int N, double *x,
                             double *w)
{
  if (N < 1)
    return -1;
  x[0] = -1;
  x[N] = 1;
  if (N > 1 && orc_jacobi_gauss(alpha + 1, beta + 1, N - 2, x + 1, w + 1))
    return -1;
  orc_jacobi_p(alpha, beta, N, N + 1, x, w);
  const double fac = (2 * N + alpha + beta + 1) / (N * (N + alpha + beta + 1));
  for (int k = 0; k < N + 1; ++k)
    w[k] = fac / (w[k] * w[k]);
  w[0] *= (1 + beta);
  w[N] *= (1 + alpha);
  return 0;
}

/* V[j*nx + i] = p_j(x_i) (3042) */
void orc_vandermonde(int N, int nx, const double *x, double *V)
{
  for (int j = 0; j <= N; ++j)
    orc_jacobi_p(0, 0, j, nx, x, V + j * nx);
}

void orc_grad_vandermonde(int N, int nx, const double *x, double *V)
{
  for (int j = 0; j <= N; ++j)
    orc_grad_jacobi_p(0, 0, j, nx, x, V + j * nx);
}

/* I = V_x / V (3063) */
void orc_interpolation(int N, int nx, const double *x, const double *V,
                       double *I)
{
  double *Vx = malloc(sizeof(double) * (N + 1) * nx);
  orc_vandermonde(N, nx, x, Vx);
  forwardslash(nx, N + 1, Vx, V, I);
  free(Vx);
}

/* D = V_r / V (3082) */
void orc_differentiation(int N, const double *r, const double *V, double *D)
{
  const int nx = N + 1;
  double *Vx = malloc(sizeof(double) * (N + 1) * nx);
  orc_grad_vandermonde(N, nx, r, Vx);
  forwardslash(nx, N + 1, Vx, V, D);
  free(Vx);
}

/* M = V^{-T} V^{-1} (3101) */
void orc_mass(int N, const double *V, double *M)
{
  const int n = N + 1;
  double *I = calloc((size_t)n * n, sizeof(double));
  double *invV = malloc(sizeof(double) * n * n);
  double *invVT = malloc(sizeof(double) * n * n);
  for (int i = 0; i < n * n; ++i)
    M[i] = 0;
  for (int i = 0; i < n; ++i)
    I[n * i + i] = 1;
  orc_backslash(n, n, V, I, invV);
  mtranspose(n, n, invV, n, invVT, n);
  mmmult(n, n, n, invVT, n, invV, n, M, n);
  free(I);
  free(invV);
  free(invVT);
}

/* per-N volume operators as stored by generic_init (7087-7155) */
int orc_dgx_ops(int N, double *r, double *w, double *wi, double *Dr)
{
  const int n = N + 1;
  double *V = malloc(sizeof(double) * n * n);
  if (orc_jacobi_gauss_lobatto(0, 0, N, r, w))
  {
    free(V);
    return -1;
  }
  orc_vandermonde(N, n, r, V);
  orc_differentiation(N, r, V, Dr);
  for (int i = 0; i < n; ++i)
    wi[i] = 1.0 / w[i];
  free(V);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* mortar operators (5424-5735)                                             */
/* ------------------------------------------------------------------------ */

/* GLL grid, Vandermonde and exact mass of order N (5459) */
static int grid_data(int N, double *r, double *w, double *V, double *M)
{
  if (orc_jacobi_gauss_lobatto(0, 0, N, r, w))
    return -1;
  orc_vandermonde(N, N + 1, r, V);
  orc_mass(N, V, M);
  return 0;
}

/* I_a2g = interpolation a -> g grid, P_g2a = V_a V_a^T I_a2g^T M_g (5471) */
static void interp_proj_data(int Na, int Ng, const double *V_a,
                             const double *M_g, const double *r_g,
                             double *I_a2g, double *P_g2a)
{
  const int na = Na + 1, ng = Ng + 1;
  double *MP = calloc((size_t)na * ng, sizeof(double));
  double *VtMP = calloc((size_t)na * ng, sizeof(double));
  orc_interpolation(Na, ng, r_g, V_a, I_a2g);
  for (int i = 0; i < na * ng; ++i)
    P_g2a[i] = 0;
  mTmmult(na, ng, ng, I_a2g, ng, M_g, ng, MP, na);
  mTmmult(na, ng, na, V_a, na, MP, na, VtMP, na);
  mmmult(na, ng, na, V_a, na, VtMP, na, P_g2a, na);
  free(MP);
  free(VtMP);
}

/* children grids at (r -+ 1)/2 with half mass (5516) */
static void hanging_data(int N, const double *r_f, const double *M,
                         const double *V, double *P_b2f, double *P_t2f,
                         double *I_f2b, double *I_f2t)
{
  const int n = N + 1;
  double *r_t = malloc(sizeof(double) * n);
  double *r_b = malloc(sizeof(double) * n);
  double *Mh = malloc(sizeof(double) * n * n);
  for (int k = 0; k < n; ++k)
  {
    r_t[k] = 0.5 * (r_f[k] + 1);
    r_b[k] = 0.5 * (r_f[k] - 1);
  }
  for (int k = 0; k < n * n; ++k)
    Mh[k] = 0.5 * M[k];
  interp_proj_data(N, N, V, Mh, r_t, I_f2t, P_t2f);
  interp_proj_data(N, N, V, Mh, r_b, I_f2b, P_b2f);
  free(r_t);
  free(r_b);
  free(Mh);
}

/* compose glue operators into a -> b and scale by masses (5543) */
static int multiply_projections(int Nb, int Na, int Ng, const double *P_g2b,
                                const double *P_g2g, const double *P_a2g,
                                double *P_a2b, const double *M_b,
                                double *MP_a2b, const double *w_b,
                                double *wiMP_a2b)
{
  const int nb = Nb + 1, ng = Ng + 1, na = Na + 1;
  if (!P_a2b)
  {
    for (int i = 0; i < na * nb; ++i)
      MP_a2b[i] = M_b[i];
    for (int j = 0; j < na; ++j)
      for (int i = 0; i < nb; ++i)
        wiMP_a2b[i + j * nb] = M_b[i + j * nb] / w_b[i];
    return 0;
  }
  double *tmp = calloc((size_t)nb * na, sizeof(double));
  double *l_MP = calloc((size_t)nb * na, sizeof(double));
  const double *l_P = tmp;
  if (!P_g2b && P_g2g && P_a2g)
    mmmult(nb, na, ng, P_g2g, nb, P_a2g, ng, tmp, nb);
  else if (P_g2b && P_g2g && !P_a2g)
    mmmult(nb, na, ng, P_g2b, nb, P_g2g, ng, tmp, nb);
  else if (!P_g2b && !P_g2g && P_a2g)
    l_P = P_a2g;
  else if (!P_g2b && P_g2g && !P_a2g)
    l_P = P_g2g;
  else if (P_g2b && !P_g2g && !P_a2g)
    l_P = P_g2b;
  else
  {
    free(tmp);
    free(l_MP);
    return -1;
  }
  mmmult(nb, na, nb, M_b, nb, l_P, nb, l_MP, nb);
  for (int i = 0; i < na * nb; ++i)
  {
    P_a2b[i] = l_P[i];
    MP_a2b[i] = l_MP[i];
  }
  for (int j = 0; j < na; ++j)
    for (int i = 0; i < nb; ++i)
      wiMP_a2b[i + j * nb] = l_MP[i + j * nb] / w_b[i];
  free(tmp);
  free(l_MP);
  return 0;
}

/* create_interpolators (5606), a -> b direction only */
int orc_interpolator(int Na, int Nb, double *prj, double *mass_prj,
                     double *wi_mass_prj, int *prj0_null)
{
  const int Ng = ORC_MAX(Na, Nb);
  const int na = Na + 1, nb = Nb + 1, ng = Ng + 1;
  const size_t blk = (size_t)na * nb;
  int rc = 0;

  double *r_a = malloc(sizeof(double) * na), *w_a = malloc(sizeof(double) * na);
  double *V_a = malloc(sizeof(double) * na * na);
  double *M_a = malloc(sizeof(double) * na * na);
  double *r_b = malloc(sizeof(double) * nb), *w_b = malloc(sizeof(double) * nb);
  double *V_b = malloc(sizeof(double) * nb * nb);
  double *M_b = malloc(sizeof(double) * nb * nb);
  double *r_g = malloc(sizeof(double) * ng), *w_g = malloc(sizeof(double) * ng);
  double *V_g = malloc(sizeof(double) * ng * ng);
  double *M_g = malloc(sizeof(double) * ng * ng);
  rc |= grid_data(Na, r_a, w_a, V_a, M_a);
  rc |= grid_data(Nb, r_b, w_b, V_b, M_b);
  rc |= grid_data(Ng, r_g, w_g, V_g, M_g);

  double *prj_g[5] = {NULL, NULL, NULL, NULL, NULL};
  for (int k = 1; k < 5; ++k)
    prj_g[k] = malloc(sizeof(double) * ng * ng);
  hanging_data(Ng, r_g, M_g, V_g, prj_g[1], prj_g[2], prj_g[3], prj_g[4]);

  double *I_a2g = NULL, *P_g2a = NULL, *I_b2g = NULL, *P_g2b = NULL;
  if (Na < Ng && Nb == Ng)
  {
    I_a2g = malloc(sizeof(double) * ng * na);
    P_g2a = malloc(sizeof(double) * ng * na);
    interp_proj_data(Na, Ng, V_a, M_g, r_g, I_a2g, P_g2a);
  }
  if (Na == Ng && Nb < Ng)
  {
    I_b2g = malloc(sizeof(double) * ng * nb);
    P_g2b = malloc(sizeof(double) * ng * nb);
    interp_proj_data(Nb, Ng, V_b, M_g, r_g, I_b2g, P_g2b);
  }

  *prj0_null = (Na == Nb);
  for (int k = 0; k < 5; ++k)
  {
    double *target = (k == 0 && Na == Nb) ? NULL : prj + k * blk;
    if (!target)
      memset(prj, 0, sizeof(double) * blk);
    rc |= multiply_projections(Nb, Na, Ng, P_g2b, prj_g[k], I_a2g, target, M_b,
                               mass_prj + k * blk, w_b, wi_mass_prj + k * blk);
  }

  free(I_a2g);
  free(P_g2a);
  free(I_b2g);
  free(P_g2b);
  for (int k = 1; k < 5; ++k)
    free(prj_g[k]);
  free(r_a), free(w_a), free(V_a), free(M_a);
  free(r_b), free(w_b), free(V_b), free(M_b);
  free(r_g), free(w_g), free(V_g), free(M_g);
  return rc ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* LSRK coefficient tables (8996-9090)                                      */
/* ------------------------------------------------------------------------ */

int orc_lsrk_table(int method, double *A, double *B, double *C)
{
  switch (method)
  {
  case ORC_LSRK_W33:
    A[0] = (double)(0.0L);
    A[1] = (double)(-5.0L / 9.0L);
    A[2] = (double)(-153.0L / 128.0L);
    B[0] = (double)(1.0L / 3.0L);
    B[1] = (double)(15.0L / 16.0L);
    B[2] = (double)(8.0L / 15.0L);
    C[0] = (double)(0.0L);
    C[1] = (double)(1.0L / 3.0L);
    C[2] = (double)(3.0L / 4.0L);
    C[3] = (double)(1.0L);
    return 3;
  case ORC_LSRK_HEUN:
    A[0] = (double)(0.0L);
    A[1] = -(double)(1.0L);
    B[0] = (double)(1.0L);
    B[1] = (double)(1.0L / 2.0L);
    C[0] = (double)(0.0L);
    C[1] = (double)(1.0L);
    C[2] = (double)(1.0L);
    return 2;
  case ORC_LSRK_FE:
    A[0] = (double)(0.0L);
    B[0] = (double)(1.0L);
    C[0] = (double)(0.0L);
    C[1] = (double)(1.0L);
    return 1;
  case ORC_LSRK_KC54:
  default: /* the reference warns and falls back to KC54 (9014) */
    A[0] = (double)(0.0L);
    A[1] = (double)(-567301805773.0L / 1357537059087.0L);
    A[2] = (double)(-2404267990393.0L / 2016746695238.0L);
    A[3] = (double)(-3550918686646.0L / 2091501179385.0L);
    A[4] = (double)(-1275806237668.0L / 842570457699.0L);
    B[0] = (double)(1432997174477.0L / 9575080441755.0L);
    B[1] = (double)(5161836677717.0L / 13612068292357.0L);
    B[2] = (double)(1720146321549.0L / 2090206949498.0L);
    B[3] = (double)(3134564353537.0L / 4481467310338.0L);
    B[4] = (double)(2277821191437.0L / 14882151754819.0L);
    C[0] = (double)(0.0L);
    C[1] = (double)(1432997174477.0L / 9575080441755.0L);
    C[2] = (double)(2526269341429.0L / 6820363962896.0L);
    C[3] = (double)(2006345519317.0L / 3224310063776.0L);
    C[4] = (double)(2802321613138.0L / 2924317926251.0L);
    C[5] = (double)(1.0L);
    return 5;
  }
}

/* ------------------------------------------------------------------------ */
/* integer maps (3630-3732, 6599-6877, 7430-7463)                           */
/* ------------------------------------------------------------------------ */

static const int8_t FToF_code[6][6] = {
    {0, 1, 1, 0, 0, 1}, {2, 0, 0, 1, 1, 0}, {2, 0, 0, 1, 1, 0},
    {0, 2, 2, 0, 0, 1}, {0, 2, 2, 0, 0, 1}, {2, 0, 0, 2, 2, 0}};
static const int8_t code_to_perm[3][4] = {
    {1, 2, 5, 6}, {0, 3, 4, 7}, {0, 4, 3, 7}};
static const int8_t perm_to_order[8][4] = {
    {0, 1, 2, 3}, {0, 2, 1, 3}, {1, 0, 3, 2}, {1, 3, 0, 2},
    {2, 0, 3, 1}, {2, 3, 0, 1}, {3, 1, 2, 0}, {3, 2, 1, 0}};
static const int8_t perm_to_order_inv[8][4] = {
    {0, 1, 2, 3}, {0, 2, 1, 3}, {1, 0, 3, 2}, {2, 0, 3, 1},
    {1, 3, 0, 2}, {2, 3, 0, 1}, {3, 1, 2, 0}, {3, 2, 1, 0}};

int orc_p8est_orientation(int f, int nf, int o)
{
  return code_to_perm[FToF_code[f][nf]][o];
}
int orc_p8est_perm_to_order(int bo, int h) { return perm_to_order[bo][h]; }
int orc_p8est_perm_to_order_inv(int bo, int h)
{
  return perm_to_order_inv[bo][h];
}

/* face masks of gmask_set (6599-6781); faces -x,+x,-y,+y(,-z,+z) */
void orc_gmask_faces(int dim, int N, int32_t *gmask)
{
  const int n1 = N + 1;
  if (dim == 2)
  {
    for (int i = 0; i < n1; ++i)
      gmask[0 * n1 + i] = i * n1;
    for (int i = 0; i < n1; ++i)
      gmask[1 * n1 + i] = (i + 1) * n1 - 1;
    for (int i = 0; i < n1; ++i)
      gmask[2 * n1 + i] = i;
    for (int i = 0; i < n1; ++i)
      gmask[3 * n1 + i] = n1 * N + i;
    return;
  }
  const int nfp = n1 * n1;
  int n;
  for (int f = 0; f < 6; ++f)
  {
    n = 0;
    const int fixed = (f % 2) ? N : 0;
    if (f / 2 == 0)
      for (int k = 0; k < n1; ++k)
        for (int j = 0; j < n1; ++j)
          gmask[f * nfp + n++] = fixed + j * n1 + k * n1 * n1;
    else if (f / 2 == 1)
      for (int k = 0; k < n1; ++k)
        for (int i = 0; i < n1; ++i)
          gmask[f * nfp + n++] = i + fixed * n1 + k * n1 * n1;
    else
      for (int j = 0; j < n1; ++j)
        for (int i = 0; i < n1; ++i)
          gmask[f * nfp + n++] = i + j * n1 + fixed * n1 * n1;
  }
}

/* buildmaps (6783-6877) */
void orc_buildmaps(int dim, int N, int32_t K, const int32_t *EToE,
                   const int8_t *EToF, int32_t *vmapM, int32_t *vmapP)
{
  const int n1 = N + 1;
  const int Nfaces = 2 * dim;
  const int Nfp = dim == 2 ? n1 : n1 * n1;
  const int Np = dim == 2 ? n1 * n1 : n1 * n1 * n1;
  int32_t *gmask = malloc(sizeof(int32_t) * Nfaces * Nfp);
  orc_gmask_faces(dim, N, gmask);

  for (int32_t k1 = 0, sk = 0; k1 < K; ++k1)
  {
    for (int f1 = 0; f1 < Nfaces; ++f1)
    {
      const int32_t k2 = EToE[Nfaces * k1 + f1];
      const int f2 = EToF[Nfaces * k1 + f1] % Nfaces;
      const int o = EToF[Nfaces * k1 + f1] / Nfaces;
      for (int n = 0; n < Nfp; ++n)
        vmapM[sk + n] = Np * k1 + gmask[f1 * Nfp + n];
      if (dim == 2)
      {
        for (int n = 0; n < Nfp; ++n)
          vmapP[sk + n] =
              Np * k2 + gmask[f2 * Nfp + (o ? Nfp - 1 - n : n)];
      }
      else
      {
        const int nr = FToF_code[f1][f2];
        const int ns = (k1 == k2 && f1 == f2) ? 0 : code_to_perm[nr][o];
        for (int j = 0, n = 0; j < n1; ++j)
          for (int i = 0; i < n1; ++i, ++n)
          {
            const int ir = n1 - (i + 1), jr = n1 - (j + 1);
            int oidx;
            switch (ns)
            {
            case 0: oidx = i + j * n1; break;
            case 1: oidx = j + i * n1; break;
            case 2: oidx = ir + j * n1; break;
            case 3: oidx = jr + i * n1; break;
            case 4: oidx = j + ir * n1; break;
            case 5: oidx = i + jr * n1; break;
            case 6: oidx = jr + ir * n1; break;
            default: oidx = ir + jr * n1; break;
            }
            vmapP[sk + n] = Np * k2 + gmask[f2 * Nfp + oidx];
          }
      }
      sk += Nfp;
    }
  }
  free(gmask);
}

/* glue mapOp (7430-7463) */
void orc_glue_mapOp(int dim_glue, int N, int32_t *mapOp)
{
  const int n1 = N + 1;
  if (dim_glue == 1)
  {
    for (int n = 0; n < n1; ++n)
    {
      mapOp[0 * n1 + n] = n;
      mapOp[1 * n1 + n] = n1 - (n + 1);
    }
    return;
  }
  const int nfp = n1 * n1;
  for (int j = 0; j < n1; ++j)
    for (int i = 0; i < n1; ++i)
    {
      const int ir = n1 - (i + 1), jr = n1 - (j + 1), n = i + j * n1;
      mapOp[0 * nfp + n] = i + j * n1;
      mapOp[1 * nfp + n] = j + i * n1;
      mapOp[2 * nfp + n] = ir + j * n1;
      mapOp[3 * nfp + n] = j + ir * n1;
      mapOp[4 * nfp + n] = jr + i * n1;
      mapOp[5 * nfp + n] = i + jr * n1;
      mapOp[6 * nfp + n] = jr + ir * n1;
      mapOp[7 * nfp + n] = ir + jr * n1;
    }
}

/* ------------------------------------------------------------------------ */
/* PCG32 (8328-8387), O'Neill 2014                                          */
/* ------------------------------------------------------------------------ */

uint32_t orc_pcg32_next(orc_pcg32_t *rng)
{
  const uint64_t old = rng->state;
  rng->state = old * 6364136223846793005ULL + (rng->inc | 1);
  const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  const uint32_t rot = (uint32_t)(old >> 59u);
  return (xs >> rot) | (xs << ((-rot) & 31));
}

void orc_pcg32_seed(orc_pcg32_t *rng, uint64_t initstate, uint64_t initseq)
{
  rng->state = 0U;
  rng->inc = (initseq << 1u) | 1u;
  orc_pcg32_next(rng);
  rng->state += initstate;
  orc_pcg32_next(rng);
}

uint32_t orc_pcg32_bounded(orc_pcg32_t *rng, uint32_t bound)
{
  const uint32_t threshold = -bound % bound;
  for (;;)
  {
    const uint32_t r = orc_pcg32_next(rng);
    if (r >= threshold)
      return r % bound;
  }
}

/* ------------------------------------------------------------------------ */
/* brick connectivity (stands in for p4est_mesh_new on a brick forest)      */
/* ------------------------------------------------------------------------ */

void orc_brick(int dim, const int *n, const int *periodic, int32_t *EToE,
               int8_t *EToF)
{
  const int nz = dim == 3 ? n[2] : 1;
  const int Nfaces = 2 * dim;
  for (int iz = 0; iz < nz; ++iz)
    for (int iy = 0; iy < n[1]; ++iy)
      for (int ix = 0; ix < n[0]; ++ix)
      {
        const int idx[3] = {ix, iy, iz};
        const int32_t e = ix + n[0] * (iy + n[1] * iz);
        for (int f = 0; f < Nfaces; ++f)
        {
          const int d = f / 2, s = (f % 2) ? 1 : -1;
          int j[3] = {ix, iy, iz};
          j[d] = idx[d] + s;
          int boundary = 0;
          if (j[d] < 0 || j[d] >= n[d])
          {
            if (periodic[d])
              j[d] = (j[d] + n[d]) % n[d];
            else
              boundary = 1;
          }
          if (boundary)
          {
            EToE[Nfaces * e + f] = e;
            EToF[Nfaces * e + f] = (int8_t)f;
          }
          else
          {
            EToE[Nfaces * e + f] = j[0] + n[0] * (j[1] + n[1] * j[2]);
            EToF[Nfaces * e + f] = (int8_t)(f ^ 1);
          }
        }
      }
}

/* ------------------------------------------------------------------------ */
/* right-hand side                                                          */
/* ------------------------------------------------------------------------ */

int orc_nfields(int physics) { return physics == ORC_ACOUSTIC2D ? 3 : 9; }

double orc_ricker(double f0, double t0, double t)
{
  const double pi = 4 * atan(1.0);
  const double a = (pi * f0 * (t - t0)) * (pi * f0 * (t - t0));
  return (1 - 2 * a) * exp(-a);
}

/* spatial pattern of manufactured field f at a point (sin/cos products) */
static double mms_pattern(int f, const double *S, const double *Cc)
{
  switch (f)
  {
  case 0: return S[0] * Cc[1] * Cc[2];
  case 1: return Cc[0] * S[1] * Cc[2];
  case 2: return Cc[0] * Cc[1] * S[2];
  case 3:
  case 4:
  case 5: return Cc[0] * Cc[1] * Cc[2];
  case 6: return S[0] * S[1] * Cc[2];
  case 7: return S[0] * Cc[1] * S[2];
  default: return Cc[0] * S[1] * S[2];
  }
}

/*
 * Manufactured elastic solution with amplitudes A[9]:
 *   v_f = A_f P_f(x) cos(w t), sigma_f = A_f P_f(x) sin(w t)
 * and F = dq/dt - L q, which is again F_f P_f(x) times sin (velocities) or
 * cos (stresses).
 */
void orc_mms_forcing_coef(const double *A, double k, double omega, double rho,
                          double lam, double mu, double *F)
{
  const double sv = A[0] + A[1] + A[2];
  F[0] = -omega * A[0] - k * (-A[3] + A[6] + A[7]) / rho;
  F[1] = -omega * A[1] - k * (A[6] - A[4] + A[8]) / rho;
  F[2] = -omega * A[2] - k * (A[7] + A[8] - A[5]) / rho;
  F[3] = omega * A[3] - k * (lam * sv + 2 * mu * A[0]);
  F[4] = omega * A[4] - k * (lam * sv + 2 * mu * A[1]);
  F[5] = omega * A[5] - k * (lam * sv + 2 * mu * A[2]);
  F[6] = omega * A[6] + mu * k * (A[0] + A[1]);
  F[7] = omega * A[7] + mu * k * (A[0] + A[2]);
  F[8] = omega * A[8] + mu * k * (A[1] + A[2]);
}

static void node_coords(const orc_problem_t *p, int dim, int32_t e, int n,
                        double *x)
{
  const int n1 = p->N + 1;
  int idx[3] = {n % n1, (n / n1) % n1, n / (n1 * n1)};
  for (int d = 0; d < dim; ++d)
    x[d] = p->x0[dim * e + d] + 0.5 * (p->r[idx[d]] + 1) * p->h[dim * e + d];
}

void orc_mms_exact(const orc_problem_t *p, const double *A, double t,
                   double *q)
{
  const int Np = (p->N + 1) * (p->N + 1) * (p->N + 1);
  const size_t KNp = (size_t)p->K * Np;
  for (int32_t e = 0; e < p->K; ++e)
    for (int n = 0; n < Np; ++n)
    {
      double x[3], S[3], Cc[3];
      node_coords(p, 3, e, n, x);
      for (int d = 0; d < 3; ++d)
      {
        S[d] = sin(p->mms_k * x[d]);
        Cc[d] = cos(p->mms_k * x[d]);
      }
      for (int f = 0; f < 9; ++f)
        q[f * KNp + (size_t)e * Np + n] =
            A[f] * mms_pattern(f, S, Cc) *
            (f < 3 ? cos(p->mms_omega * t) : sin(p->mms_omega * t));
    }
}

/* derivative of u along axis a at node (i,j,k):  sum_m D(idx_a, m) u(..m..) */
static double dr_axis(const double *Dr, const double *u, int n1, int a, int i,
                      int j, int k)
{
  double s = 0;
  const int idx = a == 0 ? i : (a == 1 ? j : k);
  for (int m = 0; m < n1; ++m)
  {
    const int node = a == 0 ? m + n1 * (j + n1 * k)
                            : (a == 1 ? i + n1 * (m + n1 * k)
                                      : i + n1 * (j + n1 * m));
    s += Dr[idx + m * n1] * u[node];
  }
  return s;
}

/*
 * Upwind fluctuations for normal n between the minus and plus states, given
 * as velocity v[3] and traction T[3] = sigma n.  Returns dv = v* - v- and
 * dT = T* - T- (Wilcox et al. 2010; exact Riemann solution, P and S parts).
 */
static void riemann(const double *n, const double *vm, const double *Tm,
                    double Zpm, double Zsm, const double *vp, const double *Tp,
                    double Zpp, double Zsp, double *dv, double *dT)
{
  const double vnm = vm[0] * n[0] + vm[1] * n[1] + vm[2] * n[2];
  const double vnp = vp[0] * n[0] + vp[1] * n[1] + vp[2] * n[2];
  const double Tnm = Tm[0] * n[0] + Tm[1] * n[1] + Tm[2] * n[2];
  const double Tnp = Tp[0] * n[0] + Tp[1] * n[1] + Tp[2] * n[2];
  const double dvn = (Zpp * (vnp - vnm) + (Tnp - Tnm)) / (Zpm + Zpp);
  for (int c = 0; c < 3; ++c)
  {
    const double vtm = vm[c] - vnm * n[c], vtp = vp[c] - vnp * n[c];
    const double Ttm = Tm[c] - Tnm * n[c], Ttp = Tp[c] - Tnp * n[c];
    const double zs = Zsm + Zsp;
    const double dvt = zs > 0 ? (Zsp * (vtp - vtm) + (Ttp - Ttm)) / zs : 0.0;
    dv[c] = dvn * n[c] + dvt;
    dT[c] = Zpm * dvn * n[c] + Zsm * dvt;
  }
}

/* strong-form lift of the flux fluctuation (dv, dT) at volume node nm:
 * rates of v get dT / rho, stresses get the constitutive image of dv (x) n */
static void lift_elastic(double *const *r, int nm, double lift, double rho,
                         double lam, double mu, const double *nrm,
                         const double *dvv, const double *dT)
{
  const double dvn = dvv[0] * nrm[0] + dvv[1] * nrm[1] + dvv[2] * nrm[2];
  r[0][nm] += lift * dT[0] / rho;
  r[1][nm] += lift * dT[1] / rho;
  r[2][nm] += lift * dT[2] / rho;
  r[3][nm] += lift * (lam * dvn + 2 * mu * dvv[0] * nrm[0]);
  r[4][nm] += lift * (lam * dvn + 2 * mu * dvv[1] * nrm[1]);
  r[5][nm] += lift * (lam * dvn + 2 * mu * dvv[2] * nrm[2]);
  r[6][nm] += lift * mu * (dvv[0] * nrm[1] + dvv[1] * nrm[0]);
  r[7][nm] += lift * mu * (dvv[0] * nrm[2] + dvv[2] * nrm[0]);
  r[8][nm] += lift * mu * (dvv[1] * nrm[2] + dvv[2] * nrm[1]);
}

static void rhs_elastic3d(const orc_problem_t *p, double t, const double *q,
                          double *rhs)
{
  const int N = p->N, n1 = N + 1, Np = n1 * n1 * n1, Nfp = n1 * n1;
  const size_t KNp = (size_t)p->K * Np;
  const double wend = p->w[0];
  double g = 0;
  if (p->src_elem >= 0)
    g = orc_ricker(p->src_f0, p->src_t0, t);
  const double tv = p->mms ? sin(p->mms_omega * t) : 0;
  const double ts = p->mms ? cos(p->mms_omega * t) : 0;

#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < p->K; ++e)
  {
    const double *u[9];
    double *r[9];
    for (int f = 0; f < 9; ++f)
    {
      u[f] = q + f * KNp + (size_t)e * Np;
      r[f] = rhs + f * KNp + (size_t)e * Np;
    }
    const double rho = p->rho[e], lam = p->lam[e], mu = p->mu[e];
    const double gx = 2 / p->h[3 * e + 0], gy = 2 / p->h[3 * e + 1],
                 gz = 2 / p->h[3 * e + 2];
    for (int k = 0; k < n1; ++k)
      for (int j = 0; j < n1; ++j)
        for (int i = 0; i < n1; ++i)
        {
          const int n = i + n1 * (j + n1 * k);
          double dv[3][3]; /* dv[c][a] = d v_c / d x_a */
          for (int c = 0; c < 3; ++c)
          {
            dv[c][0] = gx * dr_axis(p->Dr, u[c], n1, 0, i, j, k);
            dv[c][1] = gy * dr_axis(p->Dr, u[c], n1, 1, i, j, k);
            dv[c][2] = gz * dr_axis(p->Dr, u[c], n1, 2, i, j, k);
          }
          const double ds11 = gx * dr_axis(p->Dr, u[3], n1, 0, i, j, k);
          const double ds22 = gy * dr_axis(p->Dr, u[4], n1, 1, i, j, k);
          const double ds33 = gz * dr_axis(p->Dr, u[5], n1, 2, i, j, k);
          const double ds12x = gx * dr_axis(p->Dr, u[6], n1, 0, i, j, k);
          const double ds12y = gy * dr_axis(p->Dr, u[6], n1, 1, i, j, k);
          const double ds13x = gx * dr_axis(p->Dr, u[7], n1, 0, i, j, k);
          const double ds13z = gz * dr_axis(p->Dr, u[7], n1, 2, i, j, k);
          const double ds23y = gy * dr_axis(p->Dr, u[8], n1, 1, i, j, k);
          const double ds23z = gz * dr_axis(p->Dr, u[8], n1, 2, i, j, k);
          const double div = dv[0][0] + dv[1][1] + dv[2][2];
          r[0][n] = (ds11 + ds12y + ds13z) / rho;
          r[1][n] = (ds12x + ds22 + ds23z) / rho;
          r[2][n] = (ds13x + ds23y + ds33) / rho;
          r[3][n] = lam * div + 2 * mu * dv[0][0];
          r[4][n] = lam * div + 2 * mu * dv[1][1];
          r[5][n] = lam * div + 2 * mu * dv[2][2];
          r[6][n] = mu * (dv[0][1] + dv[1][0]);
          r[7][n] = mu * (dv[0][2] + dv[2][0]);
          r[8][n] = mu * (dv[1][2] + dv[2][1]);
        }

    const double Zpm = sqrt(rho * (lam + 2 * mu)), Zsm = sqrt(rho * mu);
    for (int f = 0; f < 6; ++f)
    {
      if (p->glued && p->glued[6 * (size_t)e + f])
        continue; /* flux from the mortars */
      const int a = f / 2;
      double nrm[3] = {0, 0, 0};
      nrm[a] = (f % 2) ? 1.0 : -1.0;
      const double lift = (2 / p->h[3 * e + a]) / wend;
      for (int fn = 0; fn < Nfp; ++fn)
      {
        const size_t s = ((size_t)e * 6 + f) * Nfp + fn;
        const int32_t idM = p->vmapM[s], idP = p->vmapP[s];
        const int32_t ep = idP / Np;
        const int nm = idM - e * Np;
        const double rhop = p->rho[ep], lamp = p->lam[ep], mup = p->mu[ep];
        const double Zpp = sqrt(rhop * (lamp + 2 * mup)), Zsp = sqrt(rhop * mup);
        double vm[3], vp[3], Tm[3], Tp[3], S[2][3][3];
        /* physical boundary: plus state = own trace mirrored (free surface:
         * sigma+ = -sigma-; rigid: v+ = -v-; absorbing: zero) */
        const int kind = p->bc ? p->bc[(size_t)e * 6 + f] : 0;
        const double msv = kind == 0 ? 1 : (kind == 1 ? 1 : (kind == 2 ? -1 : 0));
        const double mss = kind == 0 ? 1 : (kind == 1 ? -1 : (kind == 2 ? 1 : 0));
        for (int side = 0; side < 2; ++side)
        {
          const size_t id = side ? (size_t)idP : (size_t)idM;
          double(*Sg)[3] = S[side];
          Sg[0][0] = q[3 * KNp + id];
          Sg[1][1] = q[4 * KNp + id];
          Sg[2][2] = q[5 * KNp + id];
          Sg[0][1] = Sg[1][0] = q[6 * KNp + id];
          Sg[0][2] = Sg[2][0] = q[7 * KNp + id];
          Sg[1][2] = Sg[2][1] = q[8 * KNp + id];
          double *v = side ? vp : vm, *T = side ? Tp : Tm;
          const double fv = side ? msv : 1.0, fs = side ? mss : 1.0;
          for (int c = 0; c < 3; ++c)
          {
            v[c] = fv * q[c * KNp + id];
            T[c] = fs * (Sg[c][0] * nrm[0] + Sg[c][1] * nrm[1] + Sg[c][2] * nrm[2]);
          }
        }
        double dvv[3], dT[3];
        riemann(nrm, vm, Tm, Zpm, Zsm, vp, Tp, Zpp, Zsp, dvv, dT);
        lift_elastic(r, nm, lift, rho, lam, mu, nrm, dvv, dT);
      }
    }

    if (e == p->src_elem)
      for (int f = 0; f < 9; ++f)
        if (p->src_amp[f] != 0)
          for (int n = 0; n < Np; ++n)
            r[f][n] += p->src_wt[n] * p->src_amp[f] * g;

    if (p->mms)
      for (int n = 0; n < Np; ++n)
      {
        double x[3], S[3], Cc[3];
        node_coords(p, 3, e, n, x);
        for (int d = 0; d < 3; ++d)
        {
          S[d] = sin(p->mms_k * x[d]);
          Cc[d] = cos(p->mms_k * x[d]);
        }
        for (int f = 0; f < 9; ++f)
          r[f][n] += p->mms_F[f] * mms_pattern(f, S, Cc) * (f < 3 ? tv : ts);
      }
  }
}

static void rhs_acoustic2d(const orc_problem_t *p, double t, const double *q,
                           double *rhs)
{
  const int N = p->N, n1 = N + 1, Np = n1 * n1, Nfp = n1;
  const size_t KNp = (size_t)p->K * Np;
  const double wend = p->w[0];
  double g = 0;
  if (p->src_elem >= 0)
    g = orc_ricker(p->src_f0, p->src_t0, t);

#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < p->K; ++e)
  {
    const double *P = q + 0 * KNp + (size_t)e * Np;
    const double *U = q + 1 * KNp + (size_t)e * Np;
    const double *V = q + 2 * KNp + (size_t)e * Np;
    double *rP = rhs + 0 * KNp + (size_t)e * Np;
    double *rU = rhs + 1 * KNp + (size_t)e * Np;
    double *rV = rhs + 2 * KNp + (size_t)e * Np;
    const double rho = p->rho[e], K = p->lam[e];
    const double gx = 2 / p->h[2 * e + 0], gy = 2 / p->h[2 * e + 1];
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n1; ++i)
      {
        const int n = i + n1 * j;
        const double px = gx * dr_axis(p->Dr, P, n1, 0, i, j, 0);
        const double py = gy * dr_axis(p->Dr, P, n1, 1, i, j, 0);
        const double ux = gx * dr_axis(p->Dr, U, n1, 0, i, j, 0);
        const double vy = gy * dr_axis(p->Dr, V, n1, 1, i, j, 0);
        rP[n] = -K * (ux + vy);
        rU[n] = -px / rho;
        rV[n] = -py / rho;
      }
    const double Zm = sqrt(rho * K);
    for (int f = 0; f < 4; ++f)
    {
      const int a = f / 2;
      double nrm[2] = {0, 0};
      nrm[a] = (f % 2) ? 1.0 : -1.0;
      const double lift = (2 / p->h[2 * e + a]) / wend;
      for (int fn = 0; fn < Nfp; ++fn)
      {
        const size_t s = ((size_t)e * 4 + f) * Nfp + fn;
        const int32_t idM = p->vmapM[s], idP = p->vmapP[s];
        const int32_t ep = idP / Np;
        const int nm = idM - e * Np;
        const double Zp = sqrt(p->rho[ep] * p->lam[ep]);
        const int kind = p->bc ? p->bc[(size_t)e * 4 + f] : 0;
        const double msv = kind == 0 ? 1 : (kind == 1 ? 1 : (kind == 2 ? -1 : 0));
        const double mss = kind == 0 ? 1 : (kind == 1 ? -1 : (kind == 2 ? 1 : 0));
        const double vnm = q[1 * KNp + idM] * nrm[0] + q[2 * KNp + idM] * nrm[1];
        const double vnp =
            msv * (q[1 * KNp + idP] * nrm[0] + q[2 * KNp + idP] * nrm[1]);
        /* traction normal component T_n = -p; a free surface mirrors p */
        const double dvn =
            (Zp * (vnp - vnm) + (-mss * q[idP] + q[idM])) / (Zm + Zp);
        rU[nm] += lift * Zm * dvn * nrm[0] / rho;
        rV[nm] += lift * Zm * dvn * nrm[1] / rho;
        rP[nm] += -lift * K * dvn;
      }
    }
    if (e == p->src_elem)
      for (int f = 0; f < 3; ++f)
        if (p->src_amp[f] != 0)
          for (int n = 0; n < Np; ++n)
            rhs[f * KNp + (size_t)e * Np + n] +=
                p->src_wt[n] * p->src_amp[f] * g;
  }
}

void orc_rhs(const orc_problem_t *p, double t, const double *q, double *rhs)
{
  if (p->physics == ORC_ACOUSTIC2D)
    rhs_acoustic2d(p, t, q, rhs);
  else
    rhs_elastic3d(p, t, q, rhs);
}

void orc_lsrk_run(const orc_problem_t *p, int method, double t0, double dt,
                  int nsteps, double *q, double *dq, double *work)
{
  double A[6], B[6], C[6];
  const int ns = orc_lsrk_table(method, A, B, C);
  const int dim = p->physics == ORC_ACOUSTIC2D ? 2 : 3;
  size_t Np = 1;
  for (int d = 0; d < dim; ++d)
    Np *= (size_t)(p->N + 1);
  const size_t n = (size_t)orc_nfields(p->physics) * p->K * Np;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i)
    dq[i] = 0;
  double t = t0;
  for (int step = 0; step < nsteps; ++step)
  {
    for (int s = 0; s < ns; ++s)
    {
      orc_rhs(p, t + C[s] * dt, q, work);
      const double a = A[s], b = B[s];
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < n; ++i)
      {
        dq[i] = a * dq[i] + dt * work[i];
        q[i] = q[i] + b * dq[i];
      }
    }
    t += dt;
  }
}

double orc_energy(const orc_problem_t *p, const double *q)
{
  const int dim = p->physics == ORC_ACOUSTIC2D ? 2 : 3;
  const int n1 = p->N + 1;
  const int Np = dim == 2 ? n1 * n1 : n1 * n1 * n1;
  const size_t KNp = (size_t)p->K * Np;
  double sum = 0, comp = 0;
  for (int32_t e = 0; e < p->K; ++e)
  {
    double J = 1;
    for (int d = 0; d < dim; ++d)
      J *= 0.5 * p->h[dim * e + d];
    const double rho = p->rho[e], lam = p->lam[e], mu = p->mu ? p->mu[e] : 0;
    for (int n = 0; n < Np; ++n)
    {
      const int i = n % n1, j = (n / n1) % n1, k = n / (n1 * n1);
      double wq = p->w[i] * p->w[j];
      if (dim == 3)
        wq *= p->w[k];
      const size_t id = (size_t)e * Np + n;
      double dens;
      if (dim == 2)
      {
        const double P = q[id], U = q[KNp + id], V = q[2 * KNp + id];
        dens = rho * (U * U + V * V) + P * P / lam;
      }
      else
      {
        const double v1 = q[id], v2 = q[KNp + id], v3 = q[2 * KNp + id];
        const double s11 = q[3 * KNp + id], s22 = q[4 * KNp + id],
                     s33 = q[5 * KNp + id], s12 = q[6 * KNp + id],
                     s13 = q[7 * KNp + id], s23 = q[8 * KNp + id];
        /* sigma : C^-1 : sigma as deviatoric part dev : dev / (2 mu) plus
         * volumetric part tr^2 / (3 (3 lambda + 2 mu)); a fluid (mu = 0)
         * carries no deviatoric stress */
        const double tr = s11 + s22 + s33, m = tr / 3;
        const double dev = (s11 - m) * (s11 - m) + (s22 - m) * (s22 - m) +
                           (s33 - m) * (s33 - m) +
                           2 * (s12 * s12 + s13 * s13 + s23 * s23);
        dens = rho * (v1 * v1 + v2 * v2 + v3 * v3) + (mu > 0 ? dev / (2 * mu) : 0.0) +
               tr * tr / (3 * (3 * lam + 2 * mu));
      }
      const double y = 0.5 * J * wq * dens - comp;
      const double tt = sum + y;
      comp = (tt - sum) - y;
      sum = tt;
    }
  }
  return sum;
}

/* ------------------------------------------------------------------------ */
/* subdomains glued by mortars.  The reference builds the glue operators but
 * has no kernel that applies them; the use restated here follows their
 * wiring in bfam_subdomain_dgx_glue_init (bfam.c:7381-7414): each side's
 * trace is interpolated to the glue grid of order N_g = max(N_m, N_p) (or the
 * order a glue_order callback sets, bfam.c:7815-7817) with
 * interpolation{0,3,4} = prj{0,3,4}(N_m -> N_g) (the hanging halves of the
 * big face by refine-to-bottom / -top along each face axis), the upwind flux
 * is evaluated on the glue grid, and the result returns to each side's face
 * nodes with the mass projection prj{0,1,2}(N_g -> N_m) scaled by the inverse
 * LGL weight (wi_mass_prj, bfam.c:5543-5586), so that the LGL lift of the
 * face values reproduces the exact surface integral of the glue flux. */

typedef struct
{
  int ready;
  double prj[5 * 81];
  double wmp[5 * 81];
  int null0;
} orc_ops_pair_t;

static const double *ops_block(orc_ops_pair_t (*cache)[9], int Na, int Nb,
                               int which, int k)
{
  orc_ops_pair_t *c = &cache[Na][Nb];
  if (!c->ready)
  {
    double mp[5 * 81];
    orc_interpolator(Na, Nb, c->prj, mp, c->wmp, &c->null0);
    if (c->null0)
    {
      const int n = Na + 1;
      for (int i = 0; i < n * n; ++i)
        c->prj[i] = 0;
      for (int i = 0; i < n; ++i)
        c->prj[i + i * n] = 1;
    }
    c->ready = 1;
  }
  const size_t blk = (size_t)(Na + 1) * (Nb + 1);
  return (which ? c->wmp : c->prj) + k * blk;
}

/* out[g + ng*h] = sum_{a,b} X[g + ng*a] Y[h + ng*b] u[a + nu*b] */
static void tensor2(int ng, int nu, const double *X, const double *Y,
                    const double *u, double *out)
{
  for (int h = 0; h < ng; ++h)
    for (int g = 0; g < ng; ++g)
    {
      double acc = 0;
      for (int b = 0; b < nu; ++b)
      {
        double row = 0;
        for (int a = 0; a < nu; ++a)
          row += X[g + ng * a] * u[a + nu * b];
        acc += Y[h + ng * b] * row;
      }
      out[g + ng * h] = acc;
    }
}

static void face_trace(const orc_problem_t *p, const double *q, int32_t e,
                       int f, double sgn, const int32_t *gm, double (*tr)[81])
{
  const int n1 = p->N + 1, Np = n1 * n1 * n1, nfp = n1 * n1, A = f / 2;
  const size_t KNp = (size_t)p->K * Np;
  static const int sf[3][3] = {{3, 6, 7}, {6, 4, 8}, {7, 8, 5}};
  for (int n = 0; n < nfp; ++n)
  {
    const size_t id = (size_t)e * Np + gm[f * nfp + n];
    for (int c = 0; c < 3; ++c)
    {
      tr[c][n] = q[c * KNp + id];
      tr[3 + c][n] = sgn * q[sf[c][A] * KNp + id];
    }
  }
}

static void mortar_flux(const orc_problem_t *const *subs, const orc_mortar_t *m,
                        orc_ops_pair_t (*cache)[9], const double *const *q,
                        double *const *rhs)
{
  const orc_problem_t *Ps = subs[m->s_sub], *Pb = subs[m->b_sub];
  const int Ns = Ps->N, Nb = Pb->N;
  /* N_g = max(N_m, N_p), or the glue_order callback's choice (bfam.c:7815-7817) */
  const int Ng = m->order > 0 ? m->order : ORC_MAX(Ns, Nb);
  const int ns1 = Ns + 1, nb1 = Nb + 1, ng1 = Ng + 1;
  const int A = m->s_face / 2;
  const double sg = (m->s_face & 1) ? 1.0 : -1.0; /* n_s = sg e_A, n_b = -n_s */
  int32_t gs[6 * 81], gb[6 * 81];
  orc_gmask_faces(3, Ns, gs);
  orc_gmask_faces(3, Nb, gb);
  double ts[6][81], tb[6][81], ms[6][81], mb[6][81];
  face_trace(Ps, q[m->s_sub], m->s_elem, m->s_face, sg, gs, ts);
  face_trace(Pb, q[m->b_sub], m->b_elem, m->b_face, sg, gb, tb);
  const int qa = m->quad < 0 ? 0 : 3 + (m->quad & 1);
  const int qb = m->quad < 0 ? 0 : 3 + ((m->quad >> 1) & 1);
  const double *Is = ops_block(cache, Ns, Ng, 0, 0);
  const double *Ia = ops_block(cache, Nb, Ng, 0, qa);
  const double *Ib = ops_block(cache, Nb, Ng, 0, qb);
  for (int c = 0; c < 6; ++c)
  {
    tensor2(ng1, ns1, Is, Is, ts[c], ms[c]);
    tensor2(ng1, nb1, Ia, Ib, tb[c], mb[c]);
  }
  const int32_t es = m->s_elem, eb = m->b_elem;
  const double rs = Ps->rho[es], ls = Ps->lam[es], us = Ps->mu[es];
  const double rb = Pb->rho[eb], lb = Pb->lam[eb], ub = Pb->mu[eb];
  const double Zps = sqrt(rs * (ls + 2 * us)), Zss = sqrt(rs * us);
  const double Zpb = sqrt(rb * (lb + 2 * ub)), Zsb = sqrt(rb * ub);
  double ns_[3] = {0, 0, 0}, nb_[3] = {0, 0, 0};
  ns_[A] = sg;
  nb_[A] = -sg;
  double fs[6][81], fb[6][81];
  for (int g = 0; g < ng1 * ng1; ++g)
  {
    double vs[3], Ts[3], vb[3], Tb[3], Tsn[3], Tbn[3], dv[3], dT[3];
    for (int c = 0; c < 3; ++c)
    {
      vs[c] = ms[c][g];
      Ts[c] = ms[3 + c][g];
      vb[c] = mb[c][g];
      Tb[c] = mb[3 + c][g];
      Tsn[c] = -Ts[c];
      Tbn[c] = -Tb[c];
    }
    riemann(ns_, vs, Ts, Zps, Zss, vb, Tb, Zpb, Zsb, dv, dT);
    for (int c = 0; c < 3; ++c)
    {
      fs[c][g] = dv[c];
      fs[3 + c][g] = dT[c];
    }
    riemann(nb_, vb, Tbn, Zpb, Zsb, vs, Tsn, Zps, Zss, dv, dT);
    for (int c = 0; c < 3; ++c)
    {
      fb[c][g] = dv[c];
      fb[3 + c][g] = dT[c];
    }
  }
  const int pa = m->quad < 0 ? 0 : 1 + (m->quad & 1);
  const int pb = m->quad < 0 ? 0 : 1 + ((m->quad >> 1) & 1);
  const double *Ws = ops_block(cache, Ng, Ns, 1, 0);
  const double *Wa = ops_block(cache, Ng, Nb, 1, pa);
  const double *Wb = ops_block(cache, Ng, Nb, 1, pb);
  double Fs[6][81], Fb[6][81];
  for (int c = 0; c < 6; ++c)
  {
    /* W is (N+1) x (Ng+1): tensor2 with the roles of the sizes swapped */
    for (int j = 0; j < ns1; ++j)
      for (int i = 0; i < ns1; ++i)
      {
        double acc = 0;
        for (int h = 0; h < ng1; ++h)
          for (int g = 0; g < ng1; ++g)
            acc += Ws[i + ns1 * g] * Ws[j + ns1 * h] * fs[c][g + ng1 * h];
        Fs[c][i + ns1 * j] = acc;
      }
    for (int j = 0; j < nb1; ++j)
      for (int i = 0; i < nb1; ++i)
      {
        double acc = 0;
        for (int h = 0; h < ng1; ++h)
          for (int g = 0; g < ng1; ++g)
            acc += Wa[i + nb1 * g] * Wb[j + nb1 * h] * fb[c][g + ng1 * h];
        Fb[c][i + nb1 * j] = acc;
      }
  }
  for (int side = 0; side < 2; ++side)
  {
    const orc_problem_t *P = side ? Pb : Ps;
    const int32_t e = side ? eb : es;
    const int f = side ? m->b_face : m->s_face;
    const int n1 = P->N + 1, Np = n1 * n1 * n1, nfp = n1 * n1;
    const int32_t *gm = side ? gb : gs;
    const size_t KNp = (size_t)P->K * Np;
    double *r[9];
    for (int c = 0; c < 9; ++c)
      r[c] = rhs[side ? m->b_sub : m->s_sub] + c * KNp + (size_t)e * Np;
    const double lift = (2 / P->h[3 * e + A]) / P->w[0];
    const double *nrm = side ? nb_ : ns_;
    double(*F)[81] = side ? Fb : Fs;
    for (int n = 0; n < nfp; ++n)
    {
      const double dv[3] = {F[0][n], F[1][n], F[2][n]};
      const double dT[3] = {F[3][n], F[4][n], F[5][n]};
      lift_elastic(r, gm[f * nfp + n], lift, P->rho[e], P->lam[e], P->mu[e], nrm,
                   dv, dT);
    }
  }
}

void orc_rhs_multi(const orc_problem_t *const *subs, int nsub,
                   const orc_mortar_t *m, int32_t nm, double t,
                   const double *const *q, double *const *rhs)
{
  for (int s = 0; s < nsub; ++s)
    orc_rhs(subs[s], t, q[s], rhs[s]);
  orc_ops_pair_t(*cache)[9] = calloc(9, sizeof(*cache));
  for (int32_t i = 0; i < nm; ++i)
    mortar_flux(subs, &m[i], cache, q, rhs);
  free(cache);
}

void orc_lsrk_run_multi(const orc_problem_t *const *subs, int nsub,
                        const orc_mortar_t *m, int32_t nm, int method,
                        double t0, double dt, int nsteps, double *const *q,
                        double *const *dq, double *const *work)
{
  double A[6], B[6], C[6];
  const int ns = orc_lsrk_table(method, A, B, C);
  size_t *n = malloc(sizeof(size_t) * (size_t)(nsub ? nsub : 1));
  for (int s = 0; s < nsub; ++s)
  {
    const size_t n1 = (size_t)subs[s]->N + 1;
    n[s] = (size_t)orc_nfields(subs[s]->physics) * subs[s]->K * n1 * n1 * n1;
    for (size_t i = 0; i < n[s]; ++i)
      dq[s][i] = 0;
  }
  double t = t0;
  for (int step = 0; step < nsteps; ++step)
  {
    for (int st = 0; st < ns; ++st)
    {
      orc_rhs_multi(subs, nsub, m, nm, t + C[st] * dt, (const double *const *)q, work);
      for (int s = 0; s < nsub; ++s)
      {
        double *qq = q[s], *d = dq[s];
        const double *w = work[s];
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < n[s]; ++i)
        {
          d[i] = A[st] * d[i] + dt * w[i];
          qq[i] = qq[i] + B[st] * d[i];
        }
      }
    }
    t += dt;
  }
  free(n);
}

/* ------------------------------------------------------------------------ */
/* field transfer between adapted meshes                                    */
/* ------------------------------------------------------------------------ */

/* y = P x along axis `axis` of a 3-D array x with extents n[3] (index
 * i + n0 (j + n1 k)); P is (nb x n[axis]) column-major; y has the same
 * extents except nb along `axis` */
static void apply_axis(const double *P, int nb, int axis, const int *n,
                       const double *x, double *y)
{
  int m[3] = {n[0], n[1], n[2]};
  m[axis] = nb;
  const int na = n[axis];
  for (int k = 0; k < m[2]; ++k)
    for (int j = 0; j < m[1]; ++j)
      for (int i = 0; i < m[0]; ++i)
      {
        int o[3] = {i, j, k};
        const int r = o[axis];
        double acc = 0;
        for (int c = 0; c < na; ++c)
        {
          o[axis] = c;
          acc += P[r + c * nb] * x[o[0] + n[0] * (o[1] + n[1] * o[2])];
        }
        y[i + m[0] * (j + m[1] * k)] = acc;
      }
}

/*
 * Transfer of nf fields from source subdomains to destination subdomains
 * after h-adaptation, driven by the reference's transfer maps
 * (bfam_domain_pxest_transfer_maps_t, bfam.h:1122-1135, as filled by
 * bfam_domain_pxest_transfer_maps_init, bfam.c:5205-5387) and the 1-D
 * operators it uploads per order pair (Pr, Pr_cfb, Pr_cft, Pr_r2b, Pr_r2t,
 * bfam.c:8399-8404, i.e. prj[0..4] of create_interpolators, bfam.c:5606).
 * The kernel that applies them is not in the reference (it was application
 * code); the operators fix it:
 *   same    (flag 0): Pr(N_src -> N_dst) along each axis (identity, i.e. a
 *                     copy, when the orders agree);
 *   refine  (flag 2): along axis d, Pr_r2b if bit d of the destination's
 *                     child id is 0, else Pr_r2t;
 *   coarsen (flag 1): the sum over the 8 source children of, along axis d,
 *                     Pr_cfb if bit d of the child's id is 0, else Pr_cft.
 * Fields in bfam layout: q[s][f * K_s * Np_s + e * Np_s + n].
 */
int orc_transfer(int nf, int32_t num_dst, const uint8_t *flags,
                 const int8_t *dst_chld, const int32_t *src_subd,
                 const int32_t *src_elem, const int8_t *c_chld,
                 const int32_t *c_subd, const int32_t *c_elem,
                 const int *src_N, const int32_t *src_K,
                 const double *const *src_q, const int32_t *dst_subd,
                 const int32_t *dst_elem, const int *dst_N,
                 const int32_t *dst_K, double *const *dst_q)
{
  int rc = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(| : rc)
  for (int32_t k = 0; k < num_dst; ++k)
  {
    const int Nd = dst_N[dst_subd[k]], nd = Nd + 1;
    const size_t Npd = (size_t)nd * nd * nd;
    const int nsrc = flags[k] == 1 ? 8 : 1;
    double *acc = calloc(Npd * nf, sizeof(double));
    for (int c = 0; c < nsrc; ++c)
    {
      int32_t ss, se;
      int cid;
      if (flags[k] == 1)
      {
        const int32_t ci = -1 - src_subd[k];
        ss = c_subd[8 * ci + c];
        se = c_elem[8 * ci + c];
        cid = c_chld[8 * ci + c];
      }
      else
      {
        ss = src_subd[k];
        se = src_elem[k];
        cid = dst_chld[k];
      }
      const int Ns = src_N[ss], ns = Ns + 1;
      const size_t Nps = (size_t)ns * ns * ns;
      const size_t blk = (size_t)ns * nd;
      double *prj = malloc(sizeof(double) * 5 * blk);
      double *mp = malloc(sizeof(double) * 5 * blk);
      double *wmp = malloc(sizeof(double) * 5 * blk);
      int null0 = 0;
      rc |= orc_interpolator(Ns, Nd, prj, mp, wmp, &null0) != 0;
      if (null0)
        for (int i = 0; i < nd; ++i)
          for (int j = 0; j < ns; ++j)
            prj[i + j * nd] = i == j;
      const double *op[3];
      for (int d = 0; d < 3; ++d)
      {
        const int bit = (cid >> d) & 1;
        const int kk = flags[k] == 0 ? 0 : (flags[k] == 2 ? 3 + bit : 1 + bit);
        op[d] = prj + kk * blk;
      }
      double *t0 = malloc(sizeof(double) * (size_t)nd * ns * ns);
      double *t1 = malloc(sizeof(double) * (size_t)nd * nd * ns);
      double *t2 = malloc(sizeof(double) * Npd);
      for (int f = 0; f < nf; ++f)
      {
        const double *x = src_q[ss] + (size_t)f * src_K[ss] * Nps + (size_t)se * Nps;
        int n0[3] = {ns, ns, ns};
        apply_axis(op[0], nd, 0, n0, x, t0);
        int n1[3] = {nd, ns, ns};
        apply_axis(op[1], nd, 1, n1, t0, t1);
        int n2[3] = {nd, nd, ns};
        apply_axis(op[2], nd, 2, n2, t1, t2);
        for (size_t n = 0; n < Npd; ++n)
          acc[f * Npd + n] += t2[n];
      }
      free(t0), free(t1), free(t2);
      free(prj), free(mp), free(wmp);
    }
    const int sd = dst_subd[k];
    for (int f = 0; f < nf; ++f)
      for (size_t n = 0; n < Npd; ++n)
        dst_q[sd][(size_t)f * dst_K[sd] * Npd + (size_t)dst_elem[k] * Npd + n] =
            acc[f * Npd + n];
    free(acc);
  }
  return rc ? -1 : 0;
}
