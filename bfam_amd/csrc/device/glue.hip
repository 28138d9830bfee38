/*
 * Mortar flux for glued faces (hanging 2:1 faces and conforming faces
 * between different orders), run before the stage kernel of a glued batch.
 *
 * The reference builds, per glue element, the interpolation of the minus
 * trace to the glue grid of order N_g = max(N_m, N_p) (prj{0,3,4}), the
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

/* face trace (v, sigma[.][A]) of one element face into LDS, lane = node */
template <class TS, class TC, class SD>
__device__ __forceinline__ void load_trace(const SD &sd, int32_t e, int f,
                                           TC (*tr)[GLUE_MAXFP], int lane)
{
  const int n1 = sd.N + 1, nfp = n1 * n1, np = nfp * n1, A = f >> 1;
  if (lane >= nfp)
    return;
  const int parity = (int)(cglob((const Clock *)sd.clock)->stage & 1);
  gptr<const TS> q = glob(static_cast<const TS *>(sd.q[parity])) + (size_t)e * 9 * np +
                     face_node<3>(f, n1, lane % n1, lane / n1);
#pragma unroll
  for (int c = 0; c < 3; ++c)
  {
    tr[c][lane] = (TC)q[c * np];
    tr[3 + c][lane] = (TC)q[sfield(c, A) * np];
  }
}

template <class TS, class TC>
__global__ __launch_bounds__(64) void mortar_kernel(GlueArgs a)
{
  __shared__ TC tb[6][GLUE_MAXFP], ts[6][GLUE_MAXFP], fs[6][GLUE_MAXFP],
      fb[6][GLUE_MAXFP];
  const int gi = blockIdx.x;
  if (gi >= a.ngroups)
    return;
  const int lane = threadIdx.x;
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
  load_trace<TS, TC>(B, b_elem, b_face, tb, lane);
  cptr<TC> Mb = (cptr<TC>)mat_tab<TC>(B) + (size_t)b_elem * MAT_STRIDE;
  const TC Zpb = Mb[MAT_ZP], Zsb = Mb[MAT_ZS];
  TC acc[6] = {0, 0, 0, 0, 0, 0};
  for (int m = 0; m < nm; ++m)
  {
    const int32_t s_sub = gp->s_sub[m], s_elem = gp->s_elem[m], s_off = gp->s_off[m];
    const int s_face = gp->s_face[m];
    const int quad = nm == 1 ? quad0 : m;
    const auto &S = *(cglob(a.subs) + s_sub);
    const int Ns = S.N, ns1 = Ns + 1, nfs = ns1 * ns1;
    const int Ng = Ns > Nb ? Ns : Nb, ng1 = Ng + 1;
    load_trace<TS, TC>(S, s_elem, s_face, ts, lane);
    cptr<TC> Ms = (cptr<TC>)mat_tab<TC>(S) + (size_t)s_elem * MAT_STRIDE;
    const TC Zps = Ms[MAT_ZP], Zss = Ms[MAT_ZS];
    const TC izp = TC(1) / (Zps + Zpb), izs = TC(1) / (Zss + Zsb);
    __syncthreads();
    if (lane < ng1 * ng1)
    {
      const int g = lane % ng1, h = lane / ng1;
      const int kI = quad < 0 ? 0 : 3 + (quad & 1);
      const int kJ = quad < 0 ? 0 : 3 + ((quad >> 1) & 1);
      gptr<const TC> Is = ops + glue_ops_index(0, Ns, Ng, 0);
      gptr<const TC> Ia = ops + glue_ops_index(0, Nb, Ng, kI);
      gptr<const TC> Ib = ops + glue_ops_index(0, Nb, Ng, kJ);
      TC us[6], ub[6];
#pragma unroll
      for (int c = 0; c < 6; ++c)
      {
        TC x = 0;
        for (int j = 0; j < ns1; ++j)
        {
          TC row = 0;
          for (int i = 0; i < ns1; ++i)
            row += Is[g + ng1 * i] * ts[c][i + ns1 * j];
          x += Is[h + ng1 * j] * row;
        }
        us[c] = x;
        TC y = 0;
        for (int j = 0; j < nb1; ++j)
        {
          TC row = 0;
          for (int i = 0; i < nb1; ++i)
            row += Ia[g + ng1 * i] * tb[c][i + nb1 * j];
          y += Ib[h + ng1 * j] * row;
        }
        ub[c] = y;
      }
      TC dv[3], dT[3];
      riemann_dispatch<TC>(A, sgs, us, us + 3, ub, ub + 3, Zps, Zss, Zpb, Zsb, izp, izs,
                           dv, dT);
#pragma unroll
      for (int c = 0; c < 3; ++c)
      {
        fs[c][lane] = dv[c];
        fs[3 + c][lane] = dT[c];
      }
      riemann_dispatch<TC>(A, -sgs, ub, ub + 3, us, us + 3, Zpb, Zsb, Zps, Zss, izp, izs,
                           dv, dT);
#pragma unroll
      for (int c = 0; c < 3; ++c)
      {
        fb[c][lane] = dv[c];
        fb[3 + c][lane] = dT[c];
      }
    }
    __syncthreads();
    if (lane < nfs)
    {
      const int i = lane % ns1, j = lane / ns1;
      gptr<const TC> Ws = ops + glue_ops_index(1, Ng, Ns, 0);
#pragma unroll
      for (int c = 0; c < 6; ++c)
      {
        TC x = 0;
        for (int h = 0; h < ng1; ++h)
        {
          TC row = 0;
          for (int g = 0; g < ng1; ++g)
            row += Ws[i + ns1 * g] * fs[c][g + ng1 * h];
          x += Ws[j + ns1 * h] * row;
        }
        buf[s_off + c * nfs + lane] = x;
      }
    }
    if (lane < nfb)
    {
      const int i = lane % nb1, j = lane / nb1;
      const int kA = quad < 0 ? 0 : 1 + (quad & 1);
      const int kB = quad < 0 ? 0 : 1 + ((quad >> 1) & 1);
      gptr<const TC> Wa = ops + glue_ops_index(1, Ng, Nb, kA);
      gptr<const TC> Wb = ops + glue_ops_index(1, Ng, Nb, kB);
#pragma unroll
      for (int c = 0; c < 6; ++c)
      {
        TC x = 0;
        for (int h = 0; h < ng1; ++h)
        {
          TC row = 0;
          for (int g = 0; g < ng1; ++g)
            row += Wa[i + nb1 * g] * fb[c][g + ng1 * h];
          x += Wb[j + nb1 * h] * row;
        }
        acc[c] += x;
      }
    }
    __syncthreads();
  }
  if (lane < nfb)
  {
#pragma unroll
    for (int c = 0; c < 6; ++c)
      buf[b_off + c * nfb + lane] = acc[c];
  }
}

hipError_t launch_mortar(int prec, const GlueArgs &a, hipStream_t s)
{
  if (a.ngroups <= 0)
    return hipSuccess;
  switch (prec)
  {
  case PREC_FP64:
    hipLaunchKernelGGL((mortar_kernel<double, double>), dim3(a.ngroups), dim3(64), 0, s, a);
    break;
  case PREC_FP32_RHS:
    hipLaunchKernelGGL((mortar_kernel<double, float>), dim3(a.ngroups), dim3(64), 0, s, a);
    break;
  case PREC_FP32:
    hipLaunchKernelGGL((mortar_kernel<float, float>), dim3(a.ngroups), dim3(64), 0, s, a);
    break;
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

} // namespace bfamx
