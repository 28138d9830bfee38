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

/* Face data and 1-D operators of the mortar kernel live in LDS padded to
 * order 7: node (a, b) of a face at a + 8 b, operator entry (row, col) at
 * row + 8 col, zero outside the actual size.  Every contraction then runs a
 * fixed eight terms, fully unrolled; the padded terms add exact zeros after
 * the actual ones, so the sums are those of the actual size. */
constexpr int GP = 8;

/* face trace (v, sigma[.][A]) of one element face into LDS, lane = padded
 * node */
template <class TS, class TC, class SD>
__device__ __forceinline__ void load_trace(const SD &sd, int32_t e, int f,
                                           TC (*tr)[GLUE_MAXFP], int lane)
{
  const int n1 = sd.N + 1, np = n1 * n1 * n1, A = f >> 1;
  const int a = lane % GP, b = lane / GP;
  if (a >= n1 || b >= n1)
  {
#pragma unroll
    for (int c = 0; c < 6; ++c)
      tr[c][lane] = TC(0);
    return;
  }
  const int parity = (int)(cglob((const Clock *)sd.clock)->stage & 1);
  gptr<const TS> q = glob(static_cast<const TS *>(sd.q[parity])) + (size_t)e * 9 * np +
                     face_node<3>(f, n1, a, b);
#pragma unroll
  for (int c = 0; c < 3; ++c)
  {
    tr[c][lane] = (TC)q[c * np];
    tr[3 + c][lane] = (TC)q[sfield(c, A) * np];
  }
}

/* the six 1-D operators of one mortar: interpolation of the s face (Is) and
 * of the b face along each axis (Ia, Ib) to the mortar grid, and the mass
 * projections back (Ws; Wa, Wb) */
enum
{
  OP_IS = 0,
  OP_IA = 1,
  OP_IB = 2,
  OP_WS = 3,
  OP_WA = 4,
  OP_WB = 5
};

/* sum_{t < 8} M[r + 8 t] x[t * sx + o], fully unrolled */
template <class TC>
__device__ __forceinline__ TC contract(const TC *M, int r, const TC *x, int o, int sx)
{
  TC acc = 0;
#pragma unroll
  for (int t = 0; t < GP; ++t)
    acc += M[r + GP * t] * x[o + sx * t];
  return acc;
}

/*
 * One workgroup per interface group, one wavefront per mortar (wave m of
 * four takes mortar m; a conforming group uses wave 0 only).  Every global
 * read of the group (the b trace, the traces of its mortars' s faces and
 * their operators) is issued up front into LDS, so the group waits on
 * memory once.  The tensor products run sum-factorised, one axis per step
 * with the partial result in LDS: lane (g, j) first contracts along i, then
 * lane (g, h) along j.  Each step sums in the same order as the nested form,
 * so the results are those of the direct double sum.  The big face's four
 * contributions are added in mortar order by wave 0, as a single wave
 * walking the mortars in turn would add them.
 */
constexpr int GLUE_WAVES = 4;

template <class TS, class TC>
__global__ __launch_bounds__(64 * GLUE_WAVES) void mortar_kernel(GlueArgs a)
{
  __shared__ TC tb[6][GLUE_MAXFP];                     /* b trace */
  __shared__ TC ts[GLUE_WAVES][6][GLUE_MAXFP];         /* s traces */
  __shared__ TC sop[GLUE_WAVES][6][GP * GP];           /* operators */
  __shared__ TC t1[GLUE_WAVES][6][GLUE_MAXFP], t2[GLUE_WAVES][6][GLUE_MAXFP];
  __shared__ TC fs[GLUE_WAVES][6][GLUE_MAXFP], fb[GLUE_WAVES][6][GLUE_MAXFP];
  const int gi = blockIdx.x;
  if (gi >= a.ngroups)
    return;
  const int lane = threadIdx.x & 63;
  /* the wave index is uniform in the wave */
  const int m = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lo = lane % GP, hi = lane / GP; /* padded (row, column) */
  /* group, descriptors, clocks and materials are read-only here: constant
   * address space, scalar loads */
  cptr<MortarGroup> gp = cglob(a.groups) + gi;
  const int32_t b_sub = gp->b_sub, b_elem = gp->b_elem, b_off = gp->b_off;
  const int b_face = gp->b_face, nm = gp->nm, quad0 = gp->quad0;
  const auto &B = *(cglob(a.subs) + b_sub);
  const int Nb = B.N, nb1 = Nb + 1, nfb = nb1 * nb1;
  const int A = b_face >> 1;
  const TC sgs = (b_face & 1) ? TC(-1) : TC(1); /* normal sign of the s side */
  gptr<const TC> ops = glob(static_cast<const TC *>(a.ops));
  gptr<TC> buf = glob(static_cast<TC *>(a.buf));
  const bool mine = m < nm; /* this wave has a mortar */

  /* ---- every load of the group ---- */
  if (m == 0)
    load_trace<TS, TC>(B, b_elem, b_face, tb, lane);
  int Ns = 0;
  if (mine)
  {
    const auto &S = *(cglob(a.subs) + gp->s_sub[m]);
    load_trace<TS, TC>(S, gp->s_elem[m], gp->s_face[m], ts[m], lane);
    Ns = S.N;
    const int Ng = gp->s_ng[m] > 0 ? gp->s_ng[m] : (Ns > Nb ? Ns : Nb);
    const int quad = nm == 1 ? quad0 : m;
    const int kI = quad < 0 ? 0 : 3 + (quad & 1), kJ = quad < 0 ? 0 : 3 + ((quad >> 1) & 1);
    const int kA = quad < 0 ? 0 : 1 + (quad & 1), kB = quad < 0 ? 0 : 1 + ((quad >> 1) & 1);
    /* block offset, rows and columns of each operator (column-major) */
    const int base[6] = {glue_ops_index(0, Ns, Ng, 0), glue_ops_index(0, Nb, Ng, kI),
                         glue_ops_index(0, Nb, Ng, kJ), glue_ops_index(1, Ng, Ns, 0),
                         glue_ops_index(1, Ng, Nb, kA), glue_ops_index(1, Ng, Nb, kB)};
    const int rows[6] = {Ng + 1, Ng + 1, Ng + 1, Ns + 1, Nb + 1, Nb + 1};
    const int cols[6] = {Ns + 1, Nb + 1, Nb + 1, Ng + 1, Ng + 1, Ng + 1};
#pragma unroll
    for (int k = 0; k < 6; ++k)
      sop[m][k][lane] = (lo < rows[k] && hi < cols[k]) ? ops[base[k] + lo + rows[k] * hi]
                                                        : TC(0);
  }
  cptr<TC> Mb = (cptr<TC>)mat_tab<TC>(B) + (size_t)b_elem * MAT_STRIDE;
  const TC Zpb = Mb[MAT_ZP], Zsb = Mb[MAT_ZS];
  __syncthreads();

  const int ns1 = Ns + 1, nfs = ns1 * ns1;
  const TC *Is = sop[m][OP_IS], *Ia = sop[m][OP_IA], *Ib = sop[m][OP_IB];
  const TC *Ws = sop[m][OP_WS], *Wa = sop[m][OP_WA], *Wb = sop[m][OP_WB];
  TC Zps = 0, Zss = 0, izp = 0, izs = 0;
  if (mine)
  {
    const auto &S = *(cglob(a.subs) + gp->s_sub[m]);
    cptr<TC> Ms = (cptr<TC>)mat_tab<TC>(S) + (size_t)gp->s_elem[m] * MAT_STRIDE;
    Zps = Ms[MAT_ZP];
    Zss = Ms[MAT_ZS];
    izp = TC(1) / (Zps + Zpb);
    /* two fluids: no shear waves, and no tangential fluctuation */
    izs = Zss + Zsb > TC(0) ? TC(1) / (Zss + Zsb) : TC(0);
    /* interpolation to the mortar grid, first axis: lane (g, j) = (lo, hi) */
#pragma unroll
    for (int c = 0; c < 6; ++c)
    {
      t1[m][c][lane] = contract(Is, lo, ts[m][c], GP * hi, 1);
      t2[m][c][lane] = contract(Ia, lo, tb[c], GP * hi, 1);
    }
  }
  __syncthreads();
  if (mine)
  {
    /* second axis and the Riemann solves: lane (g, h) = (lo, hi) */
    TC us[6], ub[6];
#pragma unroll
    for (int c = 0; c < 6; ++c)
    {
      us[c] = contract(Is, hi, t1[m][c], lo, GP);
      ub[c] = contract(Ib, hi, t2[m][c], lo, GP);
    }
    TC dv[3], dT[3];
    riemann_dispatch<TC>(A, sgs, us, us + 3, ub, ub + 3, Zps, Zss, Zpb, Zsb, izp, izs, dv,
                         dT);
#pragma unroll
    for (int c = 0; c < 3; ++c)
    {
      fs[m][c][lane] = dv[c];
      fs[m][3 + c][lane] = dT[c];
    }
    riemann_dispatch<TC>(A, -sgs, ub, ub + 3, us, us + 3, Zpb, Zsb, Zps, Zss, izp, izs, dv,
                         dT);
#pragma unroll
    for (int c = 0; c < 3; ++c)
    {
      fb[m][c][lane] = dv[c];
      fb[m][3 + c][lane] = dT[c];
    }
  }
  __syncthreads();
  if (mine)
  {
    /* mass projection back to each face, first axis: lane (i, h) = (lo, hi) */
#pragma unroll
    for (int c = 0; c < 6; ++c)
    {
      t1[m][c][lane] = contract(Ws, lo, fs[m][c], GP * hi, 1);
      t2[m][c][lane] = contract(Wa, lo, fb[m][c], GP * hi, 1);
    }
  }
  __syncthreads();
  if (mine)
  {
    /* second axis: lane (i, j) = (lo, hi) of each face; the big face's
     * contribution goes to fs[m], free now, for the sum below */
#pragma unroll
    for (int c = 0; c < 6; ++c)
    {
      const TC x = contract(Ws, hi, t1[m][c], lo, GP);
      if (lo < ns1 && hi < ns1)
        buf[gp->s_off[m] + c * nfs + lo + ns1 * hi] = x;
      fs[m][c][lane] = contract(Wb, hi, t2[m][c], lo, GP);
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
        acc += fs[k][c][lane];
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
