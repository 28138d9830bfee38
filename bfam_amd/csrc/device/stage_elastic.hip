/*
 * Fused LSRK stage for 3-D isotropic velocity-stress elasticity on affine
 * hexahedra: volume derivative + face gather + upwind flux + LGL lift +
 * 2N-storage update in one pass over HBM.
 *
 * Method (the reference ships no RHS; apps supplied it as OCCA kernels run
 * through BFAMO_KERNEL_RUN, bfam.h:1673-1719):
 *   strong-form nodal DG on tensor LGL nodes (Hesthaven & Warburton 2008),
 *   d/dx_a = (2/h_a) D applied along axis a (the BFAM_KRON_* contractions,
 *   bfam.c:3429-3626), exact Riemann (upwind) fluctuations split into P and S
 *   parts (Wilcox et al., JCP 229 (2010) 9373), lift (sJ/J)/w_0 at the LGL
 *   face nodes, and the LSRK update dq = A dq + dt L, q += B dq
 *   (bfamo_lsrk_t coefficients, bfam.c:9016-9040).
 *
 * One workgroup handles EPB elements, one thread per volume node:
 *   1. stream q[e] (NF*Np contiguous) and dq[e] from HBM; stage q in LDS
 *   2. 18 directional derivatives per node from LDS (2(N+1) flop each)
 *   3. face tasks (one thread per face node): minus trace from LDS, plus trace
 *      gathered from the neighbour's q through nbr_elem/nbr_code (the on-device
 *      replacement of the glue trace exchange), upwind fluctuations -> LDS
 *   4. node owners lift the fluctuations of their faces (fixed face order)
 *   5. dq and q_out written back (q ping-pongs so neighbours read stage input)
 */
#include "physics.hpp"

namespace bfamx
{

template <int N, int EPB, class P, int PART = PART_FUSED>
__device__ __forceinline__ void elastic_task(const SubDev &sd, const Task &tk,
                                             const StageCtx &cx, int mode,
                                             char *smem)
{
  using TS = typename P::TS;
  using TR = typename P::TR;
  using TC = typename P::TC;
  static_assert(N >= 3, "3-D face buffer aliasing needs N >= 3");
  constexpr int n1 = N + 1, Np = n1 * n1 * n1, Nfp = n1 * n1, NF = 9;
  constexpr int BS = EPB * Np;
  constexpr int NT = EPB * 6 * Nfp;
  constexpr int TPT = (NT + BS - 1) / BS;
  using SM = ElasticSmem<N, EPB, TC>;

  const int tid = threadIdx.x;
  const int le = tid / Np, n = tid - (tid / Np) * Np;
  const bool act = tid < BS && le < tk.ne;
  const int e = tk.e0 + le;
  const int i = n % n1, j = (n / n1) % n1, k = n / (n1 * n1);

  TC *sq = reinterpret_cast<TC *>(smem);
  TC *sD = sq + SM::W;
  double *smms = reinterpret_cast<double *>(smem + SM::moff);

  const TC *Dg = dr_tab<TC>(sd);
  for (int t = tid; t < n1 * n1; t += blockDim.x)
    sD[t] = Dg[t];

  const TS *qin = static_cast<const TS *>(sd.q[cx.parity]);
  TR *dqp = static_cast<TR *>(sd.dq);
  TS qs[NF];
  TR dqv[NF];
  if (act)
  {
#pragma unroll
    for (int f = 0; f < NF; ++f)
      qs[f] = qin[qidx(e, f, n, NF, Np)];
#pragma unroll
    for (int f = 0; f < NF; ++f)
      sq[(le * NF + f) * Np + n] = (TC)qs[f];
  }
  if (sd.mms)
    for (int t = tid; t < tk.ne * 6 * n1; t += blockDim.x)
      smms[t] = sd.mms_tab[(size_t)tk.e0 * 6 * n1 + t];
  __syncthreads();

  /* ---- volume: 18 directional derivatives, folded into the rates one axis
   * at a time so only six accumulators are live ---- */
  TC rate[NF];
  TC rhoi = 0, lam = 0, mu = 0;
  if (act)
  {
    const TC *M = mat_tab<TC>(sd) + (size_t)e * MAT_STRIDE;
    rhoi = M[MAT_RHOI];
    lam = M[MAT_LAM];
    mu = M[MAT_MU];
  }
  if (act && PART == PART_SURFACE)
  {
#pragma unroll
    for (int f = 0; f < NF; ++f)
      rate[f] = (TC)sd.rhs[qidx(e, f, n, NF, Np)];
  }
  if (act && PART != PART_SURFACE)
  {
    const TC *G = geo_tab<TC>(sd) + (size_t)e * GEO_STRIDE;
    const TC *u = sq + le * NF * Np;
    const TC l2m = lam + TC(2) * mu;
    {
      TC v1 = 0, v2 = 0, v3 = 0, s11 = 0, s12 = 0, s13 = 0;
#pragma unroll
      for (int m = 0; m < n1; ++m)
      {
        const TC c = sD[i + m * n1];
        const int id = m + n1 * (j + n1 * k);
        v1 += c * u[0 * Np + id];
        v2 += c * u[1 * Np + id];
        v3 += c * u[2 * Np + id];
        s11 += c * u[3 * Np + id];
        s12 += c * u[6 * Np + id];
        s13 += c * u[7 * Np + id];
      }
      const TC g = G[GEO_GX];
      v1 *= g, v2 *= g, v3 *= g, s11 *= g, s12 *= g, s13 *= g;
      rate[0] = s11;
      rate[1] = s12;
      rate[2] = s13;
      rate[3] = l2m * v1;
      rate[4] = lam * v1;
      rate[5] = lam * v1;
      rate[6] = mu * v2;
      rate[7] = mu * v3;
      rate[8] = 0;
    }
    {
      TC v1 = 0, v2 = 0, v3 = 0, s22 = 0, s12 = 0, s23 = 0;
#pragma unroll
      for (int m = 0; m < n1; ++m)
      {
        const TC c = sD[j + m * n1];
        const int id = i + n1 * (m + n1 * k);
        v1 += c * u[0 * Np + id];
        v2 += c * u[1 * Np + id];
        v3 += c * u[2 * Np + id];
        s22 += c * u[4 * Np + id];
        s12 += c * u[6 * Np + id];
        s23 += c * u[8 * Np + id];
      }
      const TC g = G[GEO_GY];
      v1 *= g, v2 *= g, v3 *= g, s22 *= g, s12 *= g, s23 *= g;
      rate[0] += s12;
      rate[1] += s22;
      rate[2] += s23;
      rate[3] += lam * v2;
      rate[4] += l2m * v2;
      rate[5] += lam * v2;
      rate[6] += mu * v1;
      rate[8] += mu * v3;
    }
    {
      TC v1 = 0, v2 = 0, v3 = 0, s33 = 0, s13 = 0, s23 = 0;
#pragma unroll
      for (int m = 0; m < n1; ++m)
      {
        const TC c = sD[k + m * n1];
        const int id = i + n1 * (j + n1 * m);
        v1 += c * u[0 * Np + id];
        v2 += c * u[1 * Np + id];
        v3 += c * u[2 * Np + id];
        s33 += c * u[5 * Np + id];
        s13 += c * u[7 * Np + id];
        s23 += c * u[8 * Np + id];
      }
      const TC g = G[GEO_GZ];
      v1 *= g, v2 *= g, v3 *= g, s33 *= g, s13 *= g, s23 *= g;
      rate[0] += s13;
      rate[1] += s23;
      rate[2] += s33;
      rate[3] += lam * v3;
      rate[4] += lam * v3;
      rate[5] += l2m * v3;
      rate[7] += mu * v1;
      rate[8] += mu * v2;
    }
    rate[0] *= rhoi;
    rate[1] *= rhoi;
    rate[2] *= rhoi;
  }

  if constexpr (PART != PART_VOLUME)
  {
  /* ---- faces: one task per face node, fluctuations kept in registers ---- */
  TC fl[TPT][6];
#pragma unroll
  for (int r = 0; r < TPT; ++r)
  {
    const int t = tid + r * BS;
    const int lt = t / (6 * Nfp);
    if (tid < BS && t < NT && lt < tk.ne)
    {
      const int rem = t - lt * 6 * Nfp;
      const int f = rem / Nfp, fn = rem - (rem / Nfp) * Nfp;
      const int et = tk.e0 + lt;
      const int a = fn % n1, b = fn / n1;
      const int nm = face_node<3>(f, n1, a, b);
      const int32_t ep = sd.nbr_elem[(size_t)et * 6 + f];
      const int code = sd.nbr_code[(size_t)et * 6 + f];
      const int oi = perm_apply(code >> 3, n1, a, b);
      const int np = face_node<3>(code & 7, n1, oi % n1, oi / n1);
      const TC *u = sq + lt * NF * Np;
      TC vm[3], sm[6], vp[3], spl[6];
#pragma unroll
      for (int c = 0; c < 3; ++c)
      {
        vm[c] = u[c * Np + nm];
        vp[c] = (TC)qin[qidx(ep, c, np, NF, Np)];
      }
#pragma unroll
      for (int c = 0; c < 6; ++c)
      {
        sm[c] = u[(3 + c) * Np + nm];
        spl[c] = (TC)qin[qidx(ep, 3 + c, np, NF, Np)];
      }
      const TC *Mm = mat_tab<TC>(sd) + (size_t)et * MAT_STRIDE;
      const TC *Mp = mat_tab<TC>(sd) + (size_t)ep * MAT_STRIDE;
      const int axis = f >> 1;
      const TC s = (f & 1) ? TC(1) : TC(-1);
      elastic_riemann<TC>(axis, s, vm, sm, Mm[MAT_ZP], Mm[MAT_ZS], vp, spl,
                          Mp[MAT_ZP], Mp[MAT_ZS], &fl[r][0], &fl[r][3]);
    }
  }
  __syncthreads(); /* every read of the staged fields is done */

  TC *sfl = sq; /* [EPB][6 faces][6 comps][Nfp] */
#pragma unroll
  for (int r = 0; r < TPT; ++r)
  {
    const int t = tid + r * BS;
    const int lt = t / (6 * Nfp);
    if (tid < BS && t < NT && lt < tk.ne)
    {
      const int rem = t - lt * 6 * Nfp;
      const int f = rem / Nfp, fn = rem - (rem / Nfp) * Nfp;
#pragma unroll
      for (int c = 0; c < 6; ++c)
        sfl[((lt * 6 + f) * 6 + c) * Nfp + fn] = fl[r][c];
    }
  }
  __syncthreads();


  /* ---- lift: faces in the fixed order -x,+x,-y,+y,-z,+z ---- */
  if (act)
  {
    const TC *G = geo_tab<TC>(sd) + (size_t)e * GEO_STRIDE;
    const TC *F = sfl + le * 36 * Nfp;
    TC flv[6];
#define BFAMX_LIFT(face, cond, fn_)                                            \
  if (cond)                                                                    \
  {                                                                            \
    _Pragma("unroll") for (int c = 0; c < 6; ++c) flv[c] =                     \
        F[((face)*6 + c) * Nfp + (fn_)];                                       \
    elastic_lift<TC>((face) >> 1, ((face)&1) ? TC(1) : TC(-1),                 \
                     G[GEO_LX + ((face) >> 1)], rhoi, lam, mu, flv, rate);     \
  }
    BFAMX_LIFT(0, i == 0, j + n1 * k)
    BFAMX_LIFT(1, i == N, j + n1 * k)
    BFAMX_LIFT(2, j == 0, i + n1 * k)
    BFAMX_LIFT(3, j == N, i + n1 * k)
    BFAMX_LIFT(4, k == 0, i + n1 * j)
    BFAMX_LIFT(5, k == N, i + n1 * j)
#undef BFAMX_LIFT
  }

  }
  if (!act)
    return;

  /* ---- point source and manufactured forcing ---- */
  if (PART != PART_SURFACE && e == sd.src_elem)
  {
    const TC g = (TC)cx.g;
    const TC w = src_tab<TC>(sd)[n];
#pragma unroll
    for (int f = 0; f < NF; ++f)
      rate[f] += w * (TC)sd.src_amp[f] * g;
  }
  if (PART != PART_SURFACE && sd.mms)
  {
    const double *T = smms + le * 6 * n1;
    const double Sv[3] = {T[0 * n1 + i], T[2 * n1 + j], T[4 * n1 + k]};
    const double Cv[3] = {T[1 * n1 + i], T[3 * n1 + j], T[5 * n1 + k]};
    const double sw = cx.sw, cw = cx.cw;
#pragma unroll
    for (int f = 0; f < NF; ++f)
      rate[f] += (TC)(sd.mms_F[f] * mms_pattern(f, Sv, Cv) * (f < 3 ? sw : cw));
  }

  /* ---- output ---- */
  if (mode == MODE_RHS || PART != PART_FUSED)
  {
#pragma unroll
    for (int f = 0; f < NF; ++f)
      sd.rhs[qidx(e, f, n, NF, Np)] = (double)rate[f];
    return;
  }
  TS *qout = static_cast<TS *>(sd.q[cx.parity ^ 1]);
#pragma unroll
  for (int f = 0; f < NF; ++f)
    dqv[f] = dqp[qidx(e, f, n, NF, Np)];
  const TR ca = (TR)cx.a, cdt = (TR)cx.dt;
  const TS cb = (TS)cx.b;
#pragma unroll
  for (int f = 0; f < NF; ++f)
  {
    const TR d = ca * dqv[f] + cdt * (TR)rate[f];
    dqp[qidx(e, f, n, NF, Np)] = d;
    qout[qidx(e, f, n, NF, Np)] = qs[f] + cb * (TS)d;
  }
}

template <int N, int EPB, class P>
__global__ __launch_bounds__(EPB *(N + 1) * (N + 1) * (N + 1)) BFAMX_WAVES void
elastic_stage_kernel(LaunchArgs a)
{
  __shared__ __align__(16) char smem[ElasticSmem<N, EPB, typename P::TC>::bytes];
  int stride, end;
  for (int ti = first_task(a, &stride, &end); ti < end; ti += stride)
  {
    const Task tk = a.tasks[ti];
    const SubDev &sd = a.subs[tk.sub];
    const StageCtx cx = stage_ctx(sd, a);
    elastic_task<N, EPB, P>(sd, tk, cx, a.mode, smem);
    __syncthreads();
    if (threadIdx.x == 0 && a.mode == MODE_STAGE)
      task_done(sd, a.sub_ntasks[tk.sub]);
  }
}

template <class T> constexpr size_t cmax(T a, T b) { return a > b ? a : b; }

/* every order 3..7 in one launch; 512 threads, elements per block by N */
template <class P>
__global__ __launch_bounds__(512) BFAMX_WAVES void elastic_stage_mixed(LaunchArgs a)
{
  using TC = typename P::TC;
  constexpr size_t B = cmax(
      cmax(cmax(ElasticSmem<3, 8, TC>::bytes, ElasticSmem<4, 4, TC>::bytes),
           cmax(ElasticSmem<5, 2, TC>::bytes, ElasticSmem<6, 1, TC>::bytes)),
      ElasticSmem<7, 1, TC>::bytes);
  __shared__ __align__(16) char smem[B];
  int stride, end;
  for (int ti = first_task(a, &stride, &end); ti < end; ti += stride)
  {
    const Task tk = a.tasks[ti];
    const SubDev &sd = a.subs[tk.sub];
    const StageCtx cx = stage_ctx(sd, a);
    switch (tk.N)
    {
    case 3: elastic_task<3, 8, P>(sd, tk, cx, a.mode, smem); break;
    case 4: elastic_task<4, 4, P>(sd, tk, cx, a.mode, smem); break;
    case 5: elastic_task<5, 2, P>(sd, tk, cx, a.mode, smem); break;
    case 6: elastic_task<6, 1, P>(sd, tk, cx, a.mode, smem); break;
    default: elastic_task<7, 1, P>(sd, tk, cx, a.mode, smem); break;
    }
    __syncthreads();
    if (threadIdx.x == 0 && a.mode == MODE_STAGE)
      task_done(sd, a.sub_ntasks[tk.sub]);
  }
}

template <class P>
static hipError_t launch_n(int N, const LaunchArgs &a, int grid, hipStream_t s)
{
  switch (N)
  {
#define BFAMX_CASE(NN, EE)                                                     \
  case NN:                                                                     \
    hipLaunchKernelGGL((elastic_stage_kernel<NN, EE, P>), dim3(grid),          \
                       dim3(EE * (NN + 1) * (NN + 1) * (NN + 1)), 0, s, a);    \
    break;
    BFAMX_CASE(3, 8)
    BFAMX_CASE(4, 4)
    BFAMX_CASE(5, 2)
    BFAMX_CASE(6, 1)
    BFAMX_CASE(7, 1)
#undef BFAMX_CASE
  default:
    hipLaunchKernelGGL((elastic_stage_mixed<P>), dim3(grid), dim3(512), 0, s,
                       a);
    break;
  }
  return hipGetLastError();
}

hipError_t launch_elastic_stage(int prec, int N, const LaunchArgs &a, int grid,
                                hipStream_t s)
{
  switch (prec)
  {
  case PREC_FP32: return launch_n<P32>(N, a, grid, s);
  case PREC_FP32_RHS: return launch_n<P32R>(N, a, grid, s);
  default: return launch_n<P64>(N, a, grid, s);
  }
}

/* unfused FP64 form: the same element bodies run as a volume kernel and a
 * surface kernel, each a separate pass over HBM through the rhs array */
template <int N, int EPB, int PART>
__global__ __launch_bounds__(EPB *(N + 1) * (N + 1) * (N + 1)) void
elastic_part_kernel(LaunchArgs a)
{
  __shared__ __align__(16) char smem[ElasticSmem<N, EPB, double>::bytes];
  int stride, end;
  for (int ti = first_task(a, &stride, &end); ti < end; ti += stride)
  {
    const Task tk = a.tasks[ti];
    const SubDev &sd = a.subs[tk.sub];
    const StageCtx cx = stage_ctx(sd, a);
    elastic_task<N, EPB, P64, PART>(sd, tk, cx, a.mode, smem);
    __syncthreads();
  }
}

hipError_t launch_elastic_plain(int N, const LaunchArgs &a, int grid,
                                hipStream_t s)
{
  switch (N)
  {
#define BFAMX_CASE(NN, EE)                                                     \
  case NN:                                                                     \
    hipLaunchKernelGGL((elastic_part_kernel<NN, EE, PART_VOLUME>), dim3(grid), \
                       dim3(EE * (NN + 1) * (NN + 1) * (NN + 1)), 0, s, a);    \
    hipLaunchKernelGGL((elastic_part_kernel<NN, EE, PART_SURFACE>),            \
                       dim3(grid), dim3(EE * (NN + 1) * (NN + 1) * (NN + 1)),  \
                       0, s, a);                                               \
    break;
    BFAMX_CASE(3, 8)
    BFAMX_CASE(4, 4)
    BFAMX_CASE(5, 2)
    BFAMX_CASE(6, 1)
    BFAMX_CASE(7, 1)
#undef BFAMX_CASE
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

} // namespace bfamx
