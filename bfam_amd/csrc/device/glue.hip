/*
 * Mortar flux for glued faces (hanging 2:1 faces and conforming faces
 * between different orders), run before the stage kernel of a glued batch.
 *
 * The reference builds, per glue element, the interpolation of the minus
 * trace to the glue grid of order N_g = max(N_m, N_p), or the order its
 * glue_order callback sets (prj{0,3,4}), the
 * projection back (prj{0,1,2}) and the mass projection
 * (bfam_subdomain_dgx_glue_init, bfam.c:7381-7414; the operators of
 * create_interpolators, bfam.c:5606-5691) but leaves applying them to the
 * application.  Here one wavefront handles one interface: the big (or
 * second) face and its 1 or 4 mortars.  Per mortar, lane (g, h) of the
 * (N_g+1)^2 mortar nodes interpolates both traces (velocity and the stress
 * column of the normal axis) with the tensor product of the 1-D operators,
 * solves the upwind Riemann problem for both sides, and the fluctuations
 * return to each face's nodes through diag(1/w) M P (wi_mass_prj): the
 * stage kernel then lifts them at the face nodes like any collocated flux,
 * which reproduces each side's exact surface integral of the mortar flux.
 * The big face sums its four mortars in registers; no atomics.
 */
#include "elastic_axis.hpp"
#include "launch.hpp"
#include "physics.hpp"

namespace bfamx
{

/* Face data of the mortar kernel live in LDS padded to order 7: node
 * (a, b) of a face at a + 8 b, zero outside the actual size; the six fields
 * of a face lie GS words apart (72 rather than 64: in a column step the
 * lanes of different fields then read different banks).  The 1-D operators come padded the same way
 * (entry (row, col) at row + 8 col, zero outside the actual size) from the
 * glue_ops8_index table.  Every contraction runs a fixed eight terms; the
 * padded terms add exact zeros after the actual ones, so the sums are those
 * of the actual size. */
constexpr int GP = 8;
constexpr int GS = 72;

/* face trace (v, sigma[.][A]) of one element face into LDS, lane = padded
 * node */
template <class TS, class TC, class SD>
__device__ __forceinline__ void load_trace(const SD &sd, int32_t e, int f, TC *tr, int lane)
{
  const int n1 = sd.N + 1, np = n1 * n1 * n1, A = f >> 1;
  const int a = lane % GP, b = lane / GP;
  if (a >= n1 || b >= n1)
  {
#pragma unroll
    for (int c = 0; c < 6; ++c)
      tr[c * GS + lane] = TC(0);
    return;
  }
  const int parity = (int)(cglob((const Clock *)sd.clock)->stage & 1);
  gptr<const TS> q = glob(static_cast<const TS *>(sd.q[parity])) + (size_t)e * 9 * np +
                     face_node<3>(f, n1, a, b);
#pragma unroll
  for (int c = 0; c < 3; ++c)
  {
    tr[c * GS + lane] = (TC)q[c * np];
    tr[(3 + c) * GS + lane] = (TC)q[sfield(c, A) * np];
  }
}

/* One line of a tensor step: the eight values X[xi + sx t] (t < 8) of a
 * padded face, times the wave-uniform operator M (entry (r, t) at r + 8 t,
 * read with scalar loads): y[r] = sum_t M[r + 8 t] x[t], each sum in
 * ascending t. */
template <class TC>
__device__ __forceinline__ void line_apply(cptr<TC> M, const TC *X, int xi, int sx, TC (&y)[GP])
{
  TC x[GP];
#pragma unroll
  for (int t = 0; t < GP; ++t)
    x[t] = X[xi + sx * t];
#pragma unroll
  for (int r = 0; r < GP; ++r)
  {
    TC acc = 0;
#pragma unroll
    for (int t = 0; t < GP; ++t)
      acc += M[r + GP * t] * x[t];
    y[r] = acc;
  }
}

/* the same line, written back in place (Y[xi + sx r] = y[r]) */
template <class TC>
__device__ __forceinline__ void line_step(cptr<TC> M, TC *X, int xi, int sx)
{
  TC y[GP];
  line_apply<TC>(M, X, xi, sx, y);
#pragma unroll
  for (int r = 0; r < GP; ++r)
    X[xi + sx * r] = y[r];
}

/* order the LDS accesses of one wavefront: the compiler keeps them on
 * their side of this point, and the hardware runs them in issue order */
__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/*
 * One workgroup per interface group, one wavefront per mortar (wave m of
 * four takes mortar m; a conforming group uses wave 0 only).  Every global
 * read of the group (the b trace and the traces of its mortars' s faces)
 * is issued up front into LDS, so the group waits on memory once; the
 * operators are read with scalar loads as they are used.
 *
 * The tensor products run sum-factorised, one axis per step, and each step
 * is a set of line tasks: lane (c, l) = (lane / 8, lane % 8), lanes 0..47,
 * applies the step's operator to line l of field c (a row of eight nodes
 * for the first axis, a column for the second) and writes the eight results
 * back over the line it read, so no lane reads what another writes within
 * a step.  The s side runs first, then the b side, so the operator is the
 * same for every lane.  Between the steps the Riemann solves run one lane
 * per mortar node.  Each sum runs in the same order as the direct double
 * sum.  The big face's four contributions are added in mortar order by
 * wave 0, as a single wave walking the mortars in turn would add them.
 *
 * Between its steps a wave reads only its own LDS, so the steps are
 * ordered by a wavefront barrier (LDS operations of one wave execute in
 * issue order); the workgroup synchronises twice, after the loads (every
 * wave reads the b trace) and before the big face's sum.
 */
constexpr int GLUE_WAVES = 4;

template <class TS, class TC>
__global__ __launch_bounds__(64 * GLUE_WAVES) void mortar_kernel(GlueArgs a)
{
  /* b trace, s traces, and per wave the s- and b-side values of each step
   * (interpolated traces, then the fluctuations, then their projections) */
  __shared__ __attribute__((aligned(16))) TC tb[6 * GS];
  __shared__ __attribute__((aligned(16))) TC ts[GLUE_WAVES][6 * GS];
  __shared__ __attribute__((aligned(16))) TC tw[GLUE_WAVES][2][6 * GS];
  const int gi = blockIdx.x;
  if (gi >= a.ngroups)
    return;
  const int lane = threadIdx.x & 63;
  /* the wave index is uniform in the wave */
  const int m = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lo = lane % GP, hi = lane / GP; /* padded (row, column) */
  /* group, descriptors, clocks, materials and operators are read-only
   * here: constant address space, scalar loads */
  cptr<MortarGroup> gp = cglob(a.groups) + gi;
  const int32_t b_sub = gp->b_sub, b_elem = gp->b_elem, b_off = gp->b_off;
  const int b_face = gp->b_face, nm = gp->nm, quad0 = gp->quad0;
  const auto &B = *(cglob(a.subs) + b_sub);
  const int Nb = B.N, nb1 = Nb + 1, nfb = nb1 * nb1;
  const int A = b_face >> 1;
  const TC sgs = (b_face & 1) ? TC(-1) : TC(1); /* normal sign of the s side */
  gptr<TC> buf = glob(static_cast<TC *>(a.buf));
  const bool mine = m < nm; /* this wave has a mortar */

  /* ---- every load of the group ---- */
  if (m == 0)
    load_trace<TS, TC>(B, b_elem, b_face, tb, lane);
  /* the mortar's orders and its six operators (glue_ops8_index blocks):
   * interpolation of the s face (Is) and of the b face along each axis
   * (Ia, Ib) to the mortar grid, and the mass projections back (Ws; Wa,
   * Wb).  Taken before the barrier below: computed after it, the waves
   * m > 0 of hanging groups were measured to use wrong operators, a
   * compiler effect found by bisection and not understood further; the
   * hanging-mortar parity tests cover it. */
  int Ns = 0, oIs = 0, oIa = 0, oIb = 0, oWs = 0, oWa = 0, oWb = 0;
  if (mine)
  {
    const auto &S = *(cglob(a.subs) + gp->s_sub[m]);
    load_trace<TS, TC>(S, gp->s_elem[m], gp->s_face[m], ts[m], lane);
    Ns = S.N;
    const int Ng = gp->s_ng[m] > 0 ? gp->s_ng[m] : (Ns > Nb ? Ns : Nb);
    const int quad = nm == 1 ? quad0 : m;
    const int kI = quad < 0 ? 0 : 3 + (quad & 1), kJ = quad < 0 ? 0 : 3 + ((quad >> 1) & 1);
    const int kA = quad < 0 ? 0 : 1 + (quad & 1), kB = quad < 0 ? 0 : 1 + ((quad >> 1) & 1);
    oIs = glue_ops8_index(0, Ns, Ng, 0);
    oIa = glue_ops8_index(0, Nb, Ng, kI);
    oIb = glue_ops8_index(0, Nb, Ng, kJ);
    oWs = glue_ops8_index(1, Ng, Ns, 0);
    oWa = glue_ops8_index(1, Ng, Nb, kA);
    oWb = glue_ops8_index(1, Ng, Nb, kB);
  }
  cptr<TC> Mb = (cptr<TC>)mat_tab<TC>(B) + (size_t)b_elem * MAT_STRIDE;
  const TC Zpb = Mb[MAT_ZP], Zsb = Mb[MAT_ZS];
  __syncthreads();

  if (mine)
  {
    const auto &S = *(cglob(a.subs) + gp->s_sub[m]);
    const int ns1 = Ns + 1, nfs = ns1 * ns1;
    cptr<TC> ops = cglob(static_cast<const TC *>(a.ops));
    cptr<TC> Is = ops + oIs, Ia = ops + oIa, Ib = ops + oIb, Ws = ops + oWs, Wa = ops + oWa,
             Wb = ops + oWb;
    cptr<TC> Ms = (cptr<TC>)mat_tab<TC>(S) + (size_t)gp->s_elem[m] * MAT_STRIDE;
    const TC Zps = Ms[MAT_ZP], Zss = Ms[MAT_ZS];
    const TC izp = TC(1) / (Zps + Zpb);
    /* two fluids: no shear waves, and no tangential fluctuation */
    const TC izs = Zss + Zsb > TC(0) ? TC(1) / (Zss + Zsb) : TC(0);
    TC *ws = tw[m][0], *wb = tw[m][1];
    const int c = hi, l = lo; /* line task: field c, line l */
    const bool task = lane < 6 * GP;

    /* interpolation to the mortar grid, first axis: row l of each face */
    if (task)
    {
      TC y[GP];
      line_apply<TC>(Is, ts[m], c * GS + GP * l, 1, y);
#pragma unroll
      for (int r = 0; r < GP; ++r)
        ws[c * GS + GP * l + r] = y[r];
      line_apply<TC>(Ia, tb, c * GS + GP * l, 1, y);
#pragma unroll
      for (int r = 0; r < GP; ++r)
        wb[c * GS + GP * l + r] = y[r];
    }
    wave_sync();
    /* second axis: column l */
    if (task)
    {
      line_step<TC>(Is, ws, c * GS + l, GP);
      line_step<TC>(Ib, wb, c * GS + l, GP);
    }
    wave_sync();

    /* the Riemann solves: lane (g, h) = (lo, hi) of the mortar grid */
    {
      TC us[6], ub[6];
#pragma unroll
      for (int f = 0; f < 6; ++f)
      {
        us[f] = ws[f * GS + lane];
        ub[f] = wb[f * GS + lane];
      }
      TC dv[3], dT[3];
      riemann_dispatch<TC>(A, sgs, us, us + 3, ub, ub + 3, Zps, Zss, Zpb, Zsb, izp, izs, dv,
                           dT);
#pragma unroll
      for (int f = 0; f < 3; ++f)
      {
        ws[f * GS + lane] = dv[f];
        ws[(3 + f) * GS + lane] = dT[f];
      }
      riemann_dispatch<TC>(A, -sgs, ub, ub + 3, us, us + 3, Zpb, Zsb, Zps, Zss, izp, izs, dv,
                           dT);
#pragma unroll
      for (int f = 0; f < 3; ++f)
      {
        wb[f * GS + lane] = dv[f];
        wb[(3 + f) * GS + lane] = dT[f];
      }
    }
    wave_sync();

    /* mass projection back to each face, first axis: row l */
    if (task)
    {
      line_step<TC>(Ws, ws, c * GS + GP * l, 1);
      line_step<TC>(Wa, wb, c * GS + GP * l, 1);
    }
    wave_sync();
    /* second axis: column l; the s face's values go to the glue buffer,
     * the big face's contribution stays in LDS for the sum below */
    if (task)
    {
      TC y[GP];
      line_apply<TC>(Ws, ws, c * GS + l, GP, y);
      if (l < ns1)
      {
#pragma unroll
        for (int r = 0; r < GP; ++r)
          if (r < ns1)
            buf[gp->s_off[m] + c * nfs + l + ns1 * r] = y[r];
      }
      line_step<TC>(Wb, wb, c * GS + l, GP);
    }
  }
  __syncthreads();
  if (m == 0 && lo < nb1 && hi < nb1)
  {
#pragma unroll
    for (int c = 0; c < 6; ++c)
    {
      TC acc = 0;
      for (int k = 0; k < nm; ++k)
        acc += tw[k][1][c * GS + lane];
      buf[b_off + c * nfb + lo + nb1 * hi] = acc;
    }
  }
}

hipError_t launch_mortar(int prec, const GlueArgs &a, hipStream_t s)
{
  if (a.ngroups <= 0)
    return hipSuccess;
  switch (prec)
  {
  case PREC_FP64:
    hipLaunchKernelGGL((mortar_kernel<double, double>), dim3(a.ngroups),
                       dim3(64 * GLUE_WAVES), 0, s, a);
    break;
  case PREC_FP32_RHS:
    hipLaunchKernelGGL((mortar_kernel<double, float>), dim3(a.ngroups),
                       dim3(64 * GLUE_WAVES), 0, s, a);
    break;
  case PREC_FP32:
  case PREC_FP32_BF16X3:
    hipLaunchKernelGGL((mortar_kernel<float, float>), dim3(a.ngroups),
                       dim3(64 * GLUE_WAVES), 0, s, a);
    break;
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

} // namespace bfamx
