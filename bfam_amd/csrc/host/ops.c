/*
 * 1-D nodal operators for tensor-product LGL elements.
 *
 * The reference builds these through orthonormal Jacobi polynomials, Newton
 * from the Hale-Townsend guess and an LU solve with complete pivoting
 * (bfam.c:2681-3135, 5424-5691).  Here they come from Kopriva, "Implementing
 * Spectral Methods for PDEs" (Springer 2009):
 *   Alg. 22/23  Legendre polynomial and Gauss nodes
 *   Alg. 24/25  q = L_{N+1} - L_{N-1} and Gauss-Lobatto nodes / weights
 *   Alg. 30/34  barycentric weights and Lagrange basis evaluation
 *   Alg. 37     derivative matrix with the negative-sum diagonal
 * so the two constructions check each other to rounding.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../bfamx_internal.h"

#define PI 3.14159265358979323846

/* L_N(x) and L_N'(x) by the three-term recurrence */
static void legendre(int N, double x, double *L, double *dL)
{
  if (N == 0)
  {
    *L = 1;
    *dL = 0;
    return;
  }
  double l2 = 1, l1 = x, d2 = 0, d1 = 1;
  for (int k = 2; k <= N; ++k)
  {
    const double l0 = ((2 * k - 1) * x * l1 - (k - 1) * l2) / k;
    const double d0 = d2 + (2 * k - 1) * l1;
    l2 = l1;
    l1 = l0;
    d2 = d1;
    d1 = d0;
  }
  *L = l1;
  *dL = d1;
}

/* q = L_{N+1} - L_{N-1}, q', L_N (Kopriva Alg. 24) */
static void q_and_L(int N, double x, double *q, double *dq, double *LN)
{
  double l2 = 1, l1 = x, d2 = 0, d1 = 1;
  for (int k = 2; k <= N; ++k)
  {
    const double l0 = ((2 * k - 1) * x * l1 - (k - 1) * l2) / k;
    const double d0 = d2 + (2 * k - 1) * l1;
    l2 = l1;
    l1 = l0;
    d2 = d1;
    d1 = d0;
  }
  const int k = N + 1;
  const double lp = ((2 * k - 1) * x * l1 - (k - 1) * l2) / k;
  const double dp = d2 + (2 * k - 1) * l1;
  *q = lp - l2;
  *dq = dp - d2;
  *LN = l1;
}

int bfamx_lgl(int N, double *r, double *w)
{
  if (N < 1 || N > 64 || !r || !w)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_lgl: N=%d out of range", N);
  if (N == 1)
  {
    r[0] = -1, r[1] = 1, w[0] = w[1] = 1;
    return BFAMX_OK;
  }
  r[0] = -1;
  r[N] = 1;
  w[0] = w[N] = 2.0 / (N * (N + 1.0));
  for (int j = 1; j <= (N + 1) / 2 - 1; ++j)
  {
    double x = -cos((j + 0.25) * PI / N - 3.0 / (8 * N * PI * (j + 0.25)));
    double q, dq, LN;
    for (int it = 0; it < 100; ++it)
    {
      q_and_L(N, x, &q, &dq, &LN);
      const double d = -q / dq;
      x += d;
      if (fabs(d) <= 4 * DBL_EPSILON * fabs(x))
        break;
    }
    q_and_L(N, x, &q, &dq, &LN);
    r[j] = x;
    r[N - j] = -x;
    w[j] = w[N - j] = 2.0 / (N * (N + 1.0) * LN * LN);
  }
  if (N % 2 == 0)
  {
    double q, dq, LN;
    q_and_L(N, 0.0, &q, &dq, &LN);
    r[N / 2] = 0;
    w[N / 2] = 2.0 / (N * (N + 1.0) * LN * LN);
  }
  return BFAMX_OK;
}

static void bary_weights(int n1, const double *x, double *lam)
{
  for (int j = 0; j < n1; ++j)
  {
    double p = 1;
    for (int k = 0; k < n1; ++k)
      if (k != j)
        p *= x[j] - x[k];
    lam[j] = 1 / p;
  }
}

int bfamx_diff_matrix(int N, const double *r, double *Dr)
{
  if (N < 1 || N > 64 || !r || !Dr)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_diff_matrix: bad arguments");
  const int n1 = N + 1;
  double lam[65];
  bary_weights(n1, r, lam);
  for (int i = 0; i < n1; ++i)
  {
    double diag = 0;
    for (int j = 0; j < n1; ++j)
      if (j != i)
      {
        const double d = (lam[j] / lam[i]) / (r[i] - r[j]);
        Dr[i + j * n1] = d;
        diag -= d;
      }
    Dr[i + i * n1] = diag;
  }
  return BFAMX_OK;
}

int bfamx_lagrange_eval(int N, const double *r, double x, double *out)
{
  if (N < 0 || N > 64 || !r || !out)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_lagrange_eval: bad arguments");
  const int n1 = N + 1;
  for (int j = 0; j < n1; ++j)
    if (x == r[j])
    {
      for (int k = 0; k < n1; ++k)
        out[k] = (k == j);
      return BFAMX_OK;
    }
  double lam[65];
  bary_weights(n1, r, lam);
  double s = 0;
  for (int j = 0; j < n1; ++j)
  {
    out[j] = lam[j] / (x - r[j]);
    s += out[j];
  }
  for (int j = 0; j < n1; ++j)
    out[j] /= s;
  return BFAMX_OK;
}

/* Gauss-Legendre nodes/weights with n points (Kopriva Alg. 23) */
static void gauss_legendre(int n, double *x, double *w)
{
  const int N = n - 1;
  if (N == 0)
  {
    x[0] = 0, w[0] = 2;
    return;
  }
  for (int j = 0; j <= (N + 1) / 2 - 1; ++j)
  {
    double xj = -cos((2 * j + 1) * PI / (2 * N + 2));
    double L, dL;
    for (int it = 0; it < 100; ++it)
    {
      legendre(N + 1, xj, &L, &dL);
      const double d = -L / dL;
      xj += d;
      if (fabs(d) <= 4 * DBL_EPSILON * fabs(xj))
        break;
    }
    legendre(N + 1, xj, &L, &dL);
    x[j] = xj;
    x[N - j] = -xj;
    w[j] = w[N - j] = 2 / ((1 - xj * xj) * dL * dL);
  }
  if (N % 2 == 0)
  {
    double L, dL;
    legendre(N + 1, 0.0, &L, &dL);
    x[N / 2] = 0;
    w[N / 2] = 2 / (dL * dL);
  }
}

/* exact mass matrix of the Lagrange basis on nodes r (order N) */
static void exact_mass(int N, const double *r, double *M)
{
  const int n1 = N + 1;
  double xg[65], wg[65], l[65 * 65];
  gauss_legendre(n1, xg, wg);
  for (int q = 0; q < n1; ++q)
    bfamx_lagrange_eval(N, r, xg[q], l + q * n1);
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n1; ++j)
    {
      double s = 0;
      for (int q = 0; q < n1; ++q)
        s += wg[q] * l[q * n1 + i] * l[q * n1 + j];
      M[i + j * n1] = s;
    }
}

/* interpolation from nodes ra (order Na) to points x (nx): I[i + j*nx] */
static void interp_matrix(int Na, const double *ra, int nx, const double *x,
                          double *I)
{
  double l[65];
  for (int i = 0; i < nx; ++i)
  {
    bfamx_lagrange_eval(Na, ra, x[i], l);
    for (int j = 0; j <= Na; ++j)
      I[i + j * nx] = l[j];
  }
}

/* X = M^{-1} B for SPD M (n x n), B n x m, all column-major (Cholesky) */
static void spd_solve(int n, const double *M, int m, const double *B,
                      double *X)
{
  double L[65 * 65];
  memset(L, 0, sizeof(L));
  for (int j = 0; j < n; ++j)
  {
    double d = M[j + j * n];
    for (int k = 0; k < j; ++k)
      d -= L[j + k * n] * L[j + k * n];
    L[j + j * n] = sqrt(d);
    for (int i = j + 1; i < n; ++i)
    {
      double s = M[i + j * n];
      for (int k = 0; k < j; ++k)
        s -= L[i + k * n] * L[j + k * n];
      L[i + j * n] = s / L[j + j * n];
    }
  }
  for (int c = 0; c < m; ++c)
  {
    double y[65];
    for (int i = 0; i < n; ++i)
    {
      double s = B[i + c * n];
      for (int k = 0; k < i; ++k)
        s -= L[i + k * n] * y[k];
      y[i] = s / L[i + i * n];
    }
    for (int i = n - 1; i >= 0; --i)
    {
      double s = y[i];
      for (int k = i + 1; k < n; ++k)
        s -= L[k + i * n] * X[k + c * n];
      X[i + c * n] = s / L[i + i * n];
    }
  }
}

/* C (m x n) = A (m x k) B (k x n), column-major */
static void matmul(int m, int n, int k, const double *A, const double *B,
                   double *Cm)
{
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i)
    {
      double s = 0;
      for (int p = 0; p < k; ++p)
        s += A[i + p * m] * B[p + j * k];
      Cm[i + j * m] = s;
    }
}

/*
 * L2 projection from points of order Nsrc (nodes rs, mass Ms scaled by
 * `scale`) onto order-Ndst nodes rd through the interpolation of the
 * destination basis at the source nodes mapped by y = c0 + c1*rs:
 *   P = M_dst^{-1} I^T (scale*M_src),   I[i,j] = l^dst_j(y_i)
 */
static void l2_projection(int Nd, const double *rd, const double *Md, int Ns,
                          const double *rs, const double *Ms, double scale,
                          double c0, double c1, double *P)
{
  const int nd = Nd + 1, ns = Ns + 1;
  double y[65], I[65 * 65], ItM[65 * 65];
  for (int i = 0; i < ns; ++i)
    y[i] = c0 + c1 * rs[i];
  interp_matrix(Nd, rd, ns, y, I); /* ns x nd */
  /* ItM (nd x ns) = I^T * (scale Ms) */
  for (int j = 0; j < ns; ++j)
    for (int i = 0; i < nd; ++i)
    {
      double s = 0;
      for (int p = 0; p < ns; ++p)
        s += I[p + i * ns] * Ms[p + j * ns];
      ItM[i + j * nd] = scale * s;
    }
  spd_solve(nd, Md, ns, ItM, P);
}

int bfamx_mortar_ops(int Na, int Nb, double *prj, double *mass_prj,
                     double *wi_mass_prj, int *prj0_identity)
{
  if (Na < 1 || Nb < 1 || Na > 32 || Nb > 32 || !prj || !mass_prj ||
      !wi_mass_prj || !prj0_identity)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_mortar_ops: bad arguments");
  const int Ng = Na > Nb ? Na : Nb;
  const int na = Na + 1, nb = Nb + 1, ng = Ng + 1;
  double ra[33], wa[33], rb[33], wb[33], rg[33], wg[33];
  bfamx_lgl(Na, ra, wa);
  bfamx_lgl(Nb, rb, wb);
  bfamx_lgl(Ng, rg, wg);
  double Mb[33 * 33], Mg[33 * 33];
  exact_mass(Nb, rb, Mb);
  exact_mass(Ng, rg, Mg);

  /* hanging operators on the glue grid:
   * 1 coarsen-from-bottom: child (lower half) -> parent, L2 projection
   * 2 coarsen-from-top, 3 refine-to-bottom (interpolate parent at children),
   * 4 refine-to-top.  Children carry half the parent's mass.            */
  double H[5][33 * 33];
  double yb[33], yt[33];
  for (int i = 0; i < ng; ++i)
  {
    yb[i] = 0.5 * (rg[i] - 1);
    yt[i] = 0.5 * (rg[i] + 1);
  }
  l2_projection(Ng, rg, Mg, Ng, rg, Mg, 0.5, -0.5, 0.5, H[1]);
  l2_projection(Ng, rg, Mg, Ng, rg, Mg, 0.5, 0.5, 0.5, H[2]);
  interp_matrix(Ng, rg, ng, yb, H[3]);
  interp_matrix(Ng, rg, ng, yt, H[4]);

  /* order change a -> g (interpolation) and g -> b (L2 projection) */
  double Ia2g[33 * 33], Pg2b[33 * 33];
  const int up = Na < Ng, down = Nb < Ng;
  if (up)
    interp_matrix(Na, ra, ng, rg, Ia2g); /* ng x na */
  if (down)
    l2_projection(Nb, rb, Mb, Ng, rg, Mg, 1.0, 0.0, 1.0, Pg2b); /* nb x ng */

  *prj0_identity = (Na == Nb);
  const size_t blk = (size_t)na * nb;
  for (int k = 0; k < 5; ++k)
  {
    double *P = prj + k * blk;
    double T[33 * 33], U[33 * 33];
    /* T = H[k] * (Ia2g or I): ng x na */
    if (k == 0)
    {
      if (up)
        memcpy(T, Ia2g, sizeof(double) * ng * na);
      else
        for (int j = 0; j < na; ++j)
          for (int i = 0; i < ng; ++i)
            T[i + j * ng] = (i == j);
    }
    else if (up)
      matmul(ng, na, ng, H[k], Ia2g, T);
    else
      memcpy(T, H[k], sizeof(double) * ng * ng);
    /* U = (Pg2b or I) * T: nb x na */
    if (down)
      matmul(nb, na, ng, Pg2b, T, U);
    else
      memcpy(U, T, sizeof(double) * nb * na);
    memcpy(P, U, sizeof(double) * blk);
    matmul(nb, na, nb, Mb, U, mass_prj + k * blk);
    for (int j = 0; j < na; ++j)
      for (int i = 0; i < nb; ++i)
        wi_mass_prj[k * blk + i + j * nb] = mass_prj[k * blk + i + j * nb] / wb[i];
  }
  return BFAMX_OK;
}
