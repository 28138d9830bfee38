/*
 * Element-wise LSRK update of the unfused FP64 path, and the FP64 reductions
 * (energy, L2 difference, error against the manufactured solution).
 *
 * Reductions produce one FP64 partial sum per element (LGL-quadrature weighted
 * J w_i w_j (w_k) sums over the element's nodes); the host adds the K
 * partials with Neumaier's compensated summation, so the result does not
 * depend on launch geometry.
 */
#include "physics.hpp"

namespace bfamx
{

/* dq = A dq + dt rhs;  q = q + B dq  over all DOF */
__global__ __launch_bounds__(256) void plain_update_kernel(const SubDev *sdp,
                                                          size_t ndof)
{
  const SubDev &sd = *(const SubDev *)glob(sdp);
  gptr<const Clock> clk = glob((const Clock *)sd.clock);
  const long long st = clk->stage;
  const int s = (int)(st % sd.nstages);
  const double a = sd.A[s], b = sd.B[s], dt = clk->dt;
  gptr<double> q = glob(static_cast<double *>(sd.q[0]));
  gptr<double> dq = glob(static_cast<double *>(sd.dq));
  gptr<const double> rhs = glob((const double *)sd.rhs);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < ndof;
       i += (size_t)gridDim.x * blockDim.x)
  {
    const double d = a * dq[i] + dt * rhs[i];
    dq[i] = d;
    q[i] = q[i] + b * d;
  }
}

/* one thread per subdomain: advance its clock past the stage just run */
__global__ void clock_advance_kernel(const SubDev *subs, int nsubs)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nsubs)
    clock_advance(*(const SubDev *)(glob(subs) + i));
}

hipError_t launch_clock_advance(const SubDev *subs, int nsubs, hipStream_t s)
{
  hipLaunchKernelGGL(clock_advance_kernel, dim3((nsubs + 63) / 64), dim3(64), 0,
                     s, subs, nsubs);
  return hipGetLastError();
}

hipError_t launch_plain_update(const SubDev *sd_dev, size_t ndof,
                               hipStream_t s, int nblocks)
{
  hipLaunchKernelGGL(plain_update_kernel, dim3(nblocks), dim3(256), 0, s,
                     sd_dev, ndof);
  return hipGetLastError();
}

/* ---------------------------------------------------------------------- */

template <class T>
__device__ __forceinline__ double block_sum(double v, double *red)
{
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1)
  {
    if ((int)threadIdx.x < w)
      red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

struct ElemQuad
{
  int dim, n1, Np, nf;
};

__device__ __forceinline__ double quad_weight(const SubDev &sd, int dim,
                                              int n1, int e, int n)
{
  const double *G = sd.geo + (size_t)e * GEO_STRIDE;
  const int i = n % n1, j = (n / n1) % n1, k = n / (n1 * n1);
  double wq = sd.w[i] * sd.w[j] * 0.5 * G[GEO_HX] * 0.5 * G[GEO_HY];
  if (dim == 3)
    wq *= sd.w[k] * 0.5 * G[GEO_HZ];
  return wq;
}

template <class T>
__device__ __forceinline__ double ld(const void *p, size_t i)
{
  return (double)static_cast<const T *>(p)[i];
}

/* sigma : C^-1 : sigma of an isotropic medium (twice the strain energy
 * density) for the stress entries f[3..8], split into its deviatoric part,
 * dev : dev / (2 mu), and its volumetric part, tr^2 / (3 (3 lambda + 2 mu)).
 * A fluid (mu = 0) carries no deviatoric stress, so its part is 0 there */
__device__ __forceinline__ double stress_energy(const double *f, double lam, double mu)
{
  const double tr = f[3] + f[4] + f[5], m = tr / 3;
  const double dev = (f[3] - m) * (f[3] - m) + (f[4] - m) * (f[4] - m) +
                     (f[5] - m) * (f[5] - m) + 2 * (f[6] * f[6] + f[7] * f[7] + f[8] * f[8]);
  return (mu > 0 ? dev / (2 * mu) : 0.0) + tr * tr / (3 * (3 * lam + 2 * mu));
}

template <class TS>
__global__ __launch_bounds__(256) void energy_kernel(SubDev sd, int parity,
                                                     double *partial)
{
  __shared__ double red[256];
  const int e = blockIdx.x;
  const int dim = sd.physics == 0 ? 2 : 3;
  const int n1 = sd.N + 1;
  const int Np = dim == 2 ? n1 * n1 : n1 * n1 * n1;
  const int NF = sd.nf;
  const double *M = sd.mat + (size_t)e * MAT_STRIDE;
  const double rho = 1.0 / M[MAT_RHOI], lam = M[MAT_LAM], mu = M[MAT_MU];
  const void *q = sd.q[parity];
  double acc = 0;
  for (int n = threadIdx.x; n < Np; n += blockDim.x)
  {
    const size_t b = (size_t)e * NF * Np + n;
    double dens;
    if (dim == 2)
    {
      const double p = ld<TS>(q, b), u = ld<TS>(q, b + Np),
                   v = ld<TS>(q, b + 2 * Np);
      dens = rho * (u * u + v * v) + p * p / lam;
    }
    else
    {
      double f[9];
      for (int c = 0; c < 9; ++c)
        f[c] = ld<TS>(q, b + (size_t)c * Np);
      dens = rho * (f[0] * f[0] + f[1] * f[1] + f[2] * f[2]) + stress_energy(f, lam, mu);
    }
    acc += 0.5 * quad_weight(sd, dim, n1, e, n) * dens;
  }
  const double r = block_sum<double>(acc, red);
  if (threadIdx.x == 0)
    partial[e] = r;
}

hipError_t launch_energy(const SubDev &sd, int state_is_fp32, int parity,
                         double *partial, hipStream_t s)
{
  if (state_is_fp32)
    hipLaunchKernelGGL(energy_kernel<float>, dim3(sd.K), dim3(256), 0, s, sd,
                       parity, partial);
  else
    hipLaunchKernelGGL(energy_kernel<double>, dim3(sd.K), dim3(256), 0, s, sd,
                       parity, partial);
  return hipGetLastError();
}

template <class TA, class TB>
__global__ __launch_bounds__(256) void l2diff_kernel(SubDev a, int pa,
                                                     SubDev b, int pb,
                                                     double *partial)
{
  __shared__ double red[256];
  const int e = blockIdx.x;
  const int dim = a.physics == 0 ? 2 : 3;
  const int n1 = a.N + 1;
  const int Np = dim == 2 ? n1 * n1 : n1 * n1 * n1;
  const int NF = a.nf;
  double d2 = 0, n2 = 0;
  for (int n = threadIdx.x; n < Np; n += blockDim.x)
  {
    const double wq = quad_weight(a, dim, n1, e, n);
    double sd2 = 0, sn2 = 0;
    for (int f = 0; f < NF; ++f)
    {
      const size_t id = ((size_t)e * NF + f) * Np + n;
      const double x = ld<TA>(a.q[pa], id), y = ld<TB>(b.q[pb], id);
      sd2 += (x - y) * (x - y);
      sn2 += y * y;
    }
    d2 += wq * sd2;
    n2 += wq * sn2;
  }
  const double r1 = block_sum<double>(d2, red);
  const double r2 = block_sum<double>(n2, red);
  if (threadIdx.x == 0)
  {
    partial[2 * (size_t)e] = r1;
    partial[2 * (size_t)e + 1] = r2;
  }
}

hipError_t launch_l2diff(const SubDev &a, int a_fp32, int pa, const SubDev &b,
                         int b_fp32, int pb, double *partial, hipStream_t s)
{
  if (a_fp32 && b_fp32)
    hipLaunchKernelGGL((l2diff_kernel<float, float>), dim3(a.K), dim3(256), 0,
                       s, a, pa, b, pb, partial);
  else if (a_fp32)
    hipLaunchKernelGGL((l2diff_kernel<float, double>), dim3(a.K), dim3(256),
                       0, s, a, pa, b, pb, partial);
  else if (b_fp32)
    hipLaunchKernelGGL((l2diff_kernel<double, float>), dim3(a.K), dim3(256),
                       0, s, a, pa, b, pb, partial);
  else
    hipLaunchKernelGGL((l2diff_kernel<double, double>), dim3(a.K), dim3(256),
                       0, s, a, pa, b, pb, partial);
  return hipGetLastError();
}

struct MmsAmp
{
  double A[9];
};

__device__ __forceinline__ double mms_pat(int f, const double *Sv,
                                          const double *Cv)
{
  switch (f)
  {
  case 0: return Sv[0] * Cv[1] * Cv[2];
  case 1: return Cv[0] * Sv[1] * Cv[2];
  case 2: return Cv[0] * Cv[1] * Sv[2];
  case 3:
  case 4:
  case 5: return Cv[0] * Cv[1] * Cv[2];
  case 6: return Sv[0] * Sv[1] * Cv[2];
  case 7: return Sv[0] * Cv[1] * Sv[2];
  default: return Cv[0] * Sv[1] * Sv[2];
  }
}

template <class TS>
__global__ __launch_bounds__(256) void mms_error_kernel(SubDev sd, int parity,
                                                        MmsAmp amp, double cwt,
                                                        double swt,
                                                        double *partial)
{
  __shared__ double red[256];
  const int e = blockIdx.x;
  const int n1 = sd.N + 1, Np = n1 * n1 * n1;
  gptr<const double> T = glob(sd.mms_tab) + (size_t)e * 6 * n1;
  const double *M = sd.mat + (size_t)e * MAT_STRIDE;
  const double rho = 1.0 / M[MAT_RHOI], lam = M[MAT_LAM], mu = M[MAT_MU];
  double d2 = 0, n2 = 0, ed = 0, en = 0;
  for (int n = threadIdx.x; n < Np; n += blockDim.x)
  {
    const int i = n % n1, j = (n / n1) % n1, k = n / (n1 * n1);
    const double Sv[3] = {T[0 * n1 + i], T[2 * n1 + j], T[4 * n1 + k]};
    const double Cv[3] = {T[1 * n1 + i], T[3 * n1 + j], T[5 * n1 + k]};
    const double wq = quad_weight(sd, 3, n1, e, n);
    double sd2 = 0, sn2 = 0, dv[9], xv[9];
    for (int f = 0; f < 9; ++f)
    {
      const double ex = amp.A[f] * mms_pat(f, Sv, Cv) * (f < 3 ? cwt : swt);
      const double x = ld<TS>(sd.q[parity], ((size_t)e * 9 + f) * Np + n);
      dv[f] = x - ex;
      xv[f] = ex;
      sd2 += (x - ex) * (x - ex);
      sn2 += ex * ex;
    }
    d2 += wq * sd2;
    n2 += wq * sn2;
    /* energy norm 0.5 (rho |v|^2 + sigma : C^-1 : sigma) of the error and of
     * the exact solution */
    for (int w2 = 0; w2 < 2; ++w2)
    {
      const double *f = w2 ? xv : dv;
      const double dens =
          rho * (f[0] * f[0] + f[1] * f[1] + f[2] * f[2]) + stress_energy(f, lam, mu);
      (w2 ? en : ed) += 0.5 * wq * dens;
    }
  }
  const double r1 = block_sum<double>(d2, red);
  const double r2 = block_sum<double>(n2, red);
  const double r3 = block_sum<double>(ed, red);
  const double r4 = block_sum<double>(en, red);
  if (threadIdx.x == 0)
  {
    partial[4 * (size_t)e] = r1;
    partial[4 * (size_t)e + 1] = r2;
    partial[4 * (size_t)e + 2] = r3;
    partial[4 * (size_t)e + 3] = r4;
  }
}

hipError_t launch_mms_error(const SubDev &sd, int state_is_fp32, int parity,
                            const double *A, double cwt, double swt,
                            double *partial, hipStream_t s)
{
  MmsAmp amp;
  for (int f = 0; f < 9; ++f)
    amp.A[f] = A[f];
  if (state_is_fp32)
    hipLaunchKernelGGL(mms_error_kernel<float>, dim3(sd.K), dim3(256), 0, s, sd,
                       parity, amp, cwt, swt, partial);
  else
    hipLaunchKernelGGL(mms_error_kernel<double>, dim3(sd.K), dim3(256), 0, s,
                       sd, parity, amp, cwt, swt, partial);
  return hipGetLastError();
}

/* per element: min over axes of h_d / c_max, c_max = cp (rho cp / rho); the
 * caller takes the minimum over elements and scales by CFL / (N+1)^2 */
__global__ __launch_bounds__(256) void cfl_kernel(SubDev sd, double *partial)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= sd.K)
    return;
  gptr<const double> G = glob(sd.geo) + (size_t)e * GEO_STRIDE;
  gptr<const double> M = glob(sd.mat) + (size_t)e * MAT_STRIDE;
  const int dim = sd.physics == 0 ? 2 : 3;
  const double c = M[MAT_ZP] * M[MAT_RHOI];
  double h = G[GEO_HX];
  if (G[GEO_HY] < h)
    h = G[GEO_HY];
  if (dim == 3 && G[GEO_HZ] < h)
    h = G[GEO_HZ];
  partial[e] = h / c;
}

hipError_t launch_cfl(const SubDev &sd, double *partial, hipStream_t s)
{
  hipLaunchKernelGGL(cfl_kernel, dim3((sd.K + 255) / 256), dim3(256), 0, s, sd,
                     partial);
  return hipGetLastError();
}

} // namespace bfamx
