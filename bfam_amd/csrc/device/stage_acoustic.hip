/*
 * Fused LSRK stage for 2-D velocity-pressure acoustics on affine
 * quadrilaterals (fields p, u, v):
 *   p_t = -K (u_x + v_y),  rho u_t = -p_x,  rho v_t = -p_y,  K = rho c^2.
 * Same structure as the elastic stage (see stage_elastic.hip); the upwind
 * fluctuation is the mu = 0 case of the elastic Riemann solution:
 *   dvn = (Z+ (vn+ - vn-) + (p- - p+)) / (Z- + Z+)
 *   rate_u += lift Z- dvn n / rho,   rate_p -= lift K dvn
 * (Hesthaven & Warburton 2008, sec. 2.4 upwind flux for linear systems).
 */
#include "physics.hpp"

namespace bfamx
{

template <int N, int EPB, class T> struct AcousticSmem
{
  static constexpr int n1 = N + 1, Np = n1 * n1;
  static constexpr int QW = EPB * 3 * Np;
  static constexpr int FW = EPB * 4 * n1;
  static constexpr int W = QW > FW ? QW : FW;
  static constexpr size_t bytes = sizeof(T) * (W + n1 * n1);
};

template <int N, int EPB, class P, int PART = PART_FUSED, class SD>
__device__ __forceinline__ void acoustic_task(const SD &sd, const Task &tk,
                                              const StageCtx &cx, int mode,
                                              char *smem)
{
  using TS = typename P::TS;
  using TR = typename P::TR;
  using TC = typename P::TC;
  constexpr int n1 = N + 1, Np = n1 * n1, Nfp = n1, NF = 3;
  constexpr int BS = EPB * Np;
  constexpr int NT = EPB * 4 * Nfp;
  constexpr int TPT = (NT + BS - 1) / BS;
  using SM = AcousticSmem<N, EPB, TC>;

  const int tid = threadIdx.x;
  const int le = tid / Np, n = tid - (tid / Np) * Np;
  const bool act = tid < BS && le < tk.ne;
  const int e = tk.e0 + le;
  const int i = n % n1, j = n / n1;

  TC *sq = reinterpret_cast<TC *>(smem);
  TC *sD = sq + SM::W;
  {
    gptr<const TC> Dg = dr_tab<TC>(sd);
    for (int t = tid; t < n1 * n1; t += blockDim.x)
      sD[t] = Dg[t];
  }

  gptr<const TS> qin = glob(static_cast<const TS *>(sd.q[cx.parity]));
  TS qs[NF];
  if (act)
  {
#pragma unroll
    for (int f = 0; f < NF; ++f)
    {
      qs[f] = qin[qidx(e, f, n, NF, Np)];
      sq[(le * NF + f) * Np + n] = (TC)qs[f];
    }
  }
  __syncthreads();

  TC rate[NF];
  TC rhoi = 0, bulk = 0, Zm = 0;
  if (act)
  {
    gptr<const TC> M = mat_tab<TC>(sd) + (size_t)e * MAT_STRIDE;
    rhoi = M[MAT_RHOI];
    bulk = M[MAT_LAM];
    Zm = M[MAT_ZP];
  }
  if (act && PART == PART_SURFACE)
  {
#pragma unroll
    for (int f = 0; f < NF; ++f)
      rate[f] = (TC)glob(sd.rhs)[qidx(e, f, n, NF, Np)];
  }
  if (act && PART != PART_SURFACE)
  {
    gptr<const TC> G = geo_tab<TC>(sd) + (size_t)e * GEO_STRIDE;
    const TC *u = sq + le * NF * Np;
    TC px = 0, ux = 0, py = 0, vy = 0;
#pragma unroll
    for (int m = 0; m < n1; ++m)
    {
      const TC c = sD[i + m * n1];
      const int id = m + n1 * j;
      px += c * u[0 * Np + id];
      ux += c * u[1 * Np + id];
    }
#pragma unroll
    for (int m = 0; m < n1; ++m)
    {
      const TC c = sD[j + m * n1];
      const int id = i + n1 * m;
      py += c * u[0 * Np + id];
      vy += c * u[2 * Np + id];
    }
    const TC gx = G[GEO_GX], gy = G[GEO_GY];
    px *= gx, ux *= gx, py *= gy, vy *= gy;
    rate[0] = -bulk * (ux + vy);
    rate[1] = -px * rhoi;
    rate[2] = -py * rhoi;
  }

  if constexpr (PART != PART_VOLUME)
  {
  TC fl[TPT];
#pragma unroll
  for (int r = 0; r < TPT; ++r)
  {
    const int t = tid + r * BS;
    const int lt = t / (4 * Nfp);
    if (tid < BS && t < NT && lt < tk.ne)
    {
      const int rem = t - lt * 4 * Nfp;
      const int f = rem / Nfp, a = rem - (rem / Nfp) * Nfp;
      const int et = tk.e0 + lt;
      const int nm = face_node<2>(f, n1, a, 0);
      const int32_t ep = glob(sd.nbr_elem)[(size_t)et * 4 + f];
      const int code = glob(sd.nbr_code)[(size_t)et * 4 + f];
      const int np = nbr_node(code, f, n1, nm);
      const TC *u = sq + lt * NF * Np;
      const int axis = f >> 1;
      const TC s = (f & 1) ? TC(1) : TC(-1);
      TC sv, ss; /* pressure plays the stress: free surface p+ = -p- */
      nbr_mirror<TC>(code, &sv, &ss);
      const TC pm = u[nm], vnm = s * u[(1 + axis) * Np + nm];
      const TC pp = ss * (TC)qin[qidx(ep, 0, np, NF, Np)];
      const TC vnp = sv * s * (TC)qin[qidx(ep, 1 + axis, np, NF, Np)];
      gptr<const TC> Mm = mat_tab<TC>(sd) + (size_t)et * MAT_STRIDE;
      gptr<const TC> Mp = mat_tab<TC>(sd) + (size_t)ep * MAT_STRIDE;
      fl[r] = (Mp[MAT_ZP] * (vnp - vnm) + (pm - pp)) / (Mm[MAT_ZP] + Mp[MAT_ZP]);
    }
  }
  __syncthreads();
  TC *sfl = sq; /* [EPB][4][Nfp] */
#pragma unroll
  for (int r = 0; r < TPT; ++r)
  {
    const int t = tid + r * BS;
    if (tid < BS && t < NT && t / (4 * Nfp) < tk.ne)
      sfl[t] = fl[r];
  }
  __syncthreads();

  if (act)
  {
    gptr<const TC> G = geo_tab<TC>(sd) + (size_t)e * GEO_STRIDE;
    const TC *F = sfl + le * 4 * Nfp;
#define BFAMX_ALIFT(face, cond, fn_)                                           \
  if (cond)                                                                    \
  {                                                                            \
    const TC lift = G[GEO_LX + ((face) >> 1)];                                 \
    const TC dvn = F[(face)*Nfp + (fn_)];                                      \
    const TC sg = ((face)&1) ? TC(1) : TC(-1);                                 \
    rate[1 + ((face) >> 1)] += lift * Zm * dvn * sg * rhoi;                    \
    rate[0] -= lift * bulk * dvn;                                              \
  }
    BFAMX_ALIFT(0, i == 0, j)
    BFAMX_ALIFT(1, i == N, j)
    BFAMX_ALIFT(2, j == 0, i)
    BFAMX_ALIFT(3, j == N, i)
#undef BFAMX_ALIFT
  }

  }
  if (!act)
    return;

  if (PART != PART_SURFACE && e == sd.src_elem)
  {
    const TC g = (TC)cx.g;
    const TC w = src_tab<TC>(sd)[n];
#pragma unroll
    for (int f = 0; f < NF; ++f)
      rate[f] += w * (TC)sd.src_amp[f] * g;
  }

  if (mode == MODE_RHS || PART != PART_FUSED)
  {
#pragma unroll
    for (int f = 0; f < NF; ++f)
      glob(sd.rhs)[qidx(e, f, n, NF, Np)] = (double)rate[f];
    return;
  }
  gptr<TS> qout = glob(static_cast<TS *>(sd.q[cx.parity ^ 1]));
  gptr<TR> dqp = glob(static_cast<TR *>(sd.dq));
  const TR ca = (TR)cx.a, cdt = (TR)cx.dt;
  const TS cb = (TS)cx.b;
  /* A = 0 (first stage of a step): no dq read.  FP32, FP64: the old dq is
   * rebuilt from the previous state, B_{s-1} dq_{s-1} = q_s - q_{s-1},
   * read from the other ping-pong buffer before this thread overwrites the
   * node (a holds A_s / B_{s-1}) */
  const bool rd_dq = cx.a != 0.0;
  constexpr bool REC = sizeof(TS) == sizeof(TR);
  const bool rec = REC && cx.rec;
#pragma unroll
  for (int f = 0; f < NF; ++f)
  {
    const size_t id = qidx(e, f, n, NF, Np);
    TR dqo = 0;
    if (rd_dq)
    {
      if constexpr (REC)
        dqo = rec ? (TR)(qs[f] - qout[id]) : dqp[id];
      else
        dqo = dqp[id];
    }
    const TR d = ca * dqo + cdt * (TR)rate[f];
    if (!cx.dq_dead && !rec)
      dqp[id] = d;
    qout[id] = qs[f] + cb * (TS)d;
  }
}

template <int N, int EPB, class P>
__global__ __launch_bounds__(EPB *(N + 1) * (N + 1)) void
acoustic_stage_kernel(LaunchArgs a)
{
  __shared__ __align__(16) char smem[AcousticSmem<N, EPB, typename P::TC>::bytes];
  int stride, end;
  for (int ti = first_task(a, &stride, &end); ti < end; ti += stride)
  {
    const Task tk = load_task(a.tasks, ti);
    const auto &sd = *(cglob(a.subs) + tk.sub);
    const StageCtx cx = stage_ctx(sd, a);
    acoustic_task<N, EPB, P>(sd, tk, cx, a.mode, smem);
    __syncthreads();
  }
}

template <class P>
static hipError_t launch_an(int N, const LaunchArgs &a, int grid, hipStream_t s)
{
  switch (N)
  {
#define BFAMX_CASE(NN)                                                         \
  case NN:                                                                     \
  {                                                                            \
    constexpr int E = 256 / ((NN + 1) * (NN + 1));                             \
    hipLaunchKernelGGL((acoustic_stage_kernel<NN, E, P>), dim3(grid),          \
                       dim3(E * (NN + 1) * (NN + 1)), 0, s, a);                \
    break;                                                                     \
  }
    BFAMX_CASE(1)
    BFAMX_CASE(2)
    BFAMX_CASE(3)
    BFAMX_CASE(4)
    BFAMX_CASE(5)
    BFAMX_CASE(6)
    BFAMX_CASE(7)
    BFAMX_CASE(8)
#undef BFAMX_CASE
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_acoustic_stage(int prec, int N, const LaunchArgs &a,
                                 int grid, hipStream_t s)
{
  switch (prec)
  {
  case PREC_FP32: return launch_an<P32>(N, a, grid, s);
  case PREC_FP32_RHS: return launch_an<P32R>(N, a, grid, s);
  default: return launch_an<P64>(N, a, grid, s);
  }
}

template <int N, int EPB, int PART>
__global__ __launch_bounds__(EPB *(N + 1) * (N + 1)) void
acoustic_part_kernel(LaunchArgs a)
{
  __shared__ __align__(16) char smem[AcousticSmem<N, EPB, double>::bytes];
  int stride, end;
  for (int ti = first_task(a, &stride, &end); ti < end; ti += stride)
  {
    const Task tk = load_task(a.tasks, ti);
    const auto &sd = *(cglob(a.subs) + tk.sub);
    const StageCtx cx = stage_ctx(sd, a);
    acoustic_task<N, EPB, P64, PART>(sd, tk, cx, a.mode, smem);
    __syncthreads();
  }
}

hipError_t launch_acoustic_plain(int N, const LaunchArgs &a, int grid,
                                 hipStream_t s)
{
  switch (N)
  {
#define BFAMX_CASE(NN)                                                         \
  case NN:                                                                     \
  {                                                                            \
    constexpr int E = 256 / ((NN + 1) * (NN + 1));                             \
    hipLaunchKernelGGL((acoustic_part_kernel<NN, E, PART_VOLUME>), dim3(grid), \
                       dim3(E * (NN + 1) * (NN + 1)), 0, s, a);                \
    hipLaunchKernelGGL((acoustic_part_kernel<NN, E, PART_SURFACE>),            \
                       dim3(grid), dim3(E * (NN + 1) * (NN + 1)), 0, s, a);    \
    break;                                                                     \
  }
    BFAMX_CASE(1)
    BFAMX_CASE(2)
    BFAMX_CASE(3)
    BFAMX_CASE(4)
    BFAMX_CASE(5)
    BFAMX_CASE(6)
    BFAMX_CASE(7)
    BFAMX_CASE(8)
#undef BFAMX_CASE
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

} // namespace bfamx
