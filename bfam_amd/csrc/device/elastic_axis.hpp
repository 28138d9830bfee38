/*
 * Axis-aligned faces of the isotropic elastic velocity-stress system: field
 * index helpers and the exact-Riemann upwind fluctuations for a face normal
 * to axis A (Wilcox, Stadler, Burstedde, Ghattas, J. Comput. Phys. 229
 * (2010) 9373, section 3.2: P and S waves split by the impedances
 * Zp = rho cp and Zs = rho cs of both sides).
 */
#ifndef BFAMX_ELASTIC_AXIS_HPP
#define BFAMX_ELASTIC_AXIS_HPP

#include "common.hpp"

namespace bfamx
{

/* stress field index of sigma[c][axis] (field order v1 v2 v3 s11 s22 s33 s12
 * s13 s23) */
__device__ __forceinline__ constexpr int sfield(int c, int a)
{
  return (c == a) ? 3 + c
                  : ((c + a == 1) ? 6 : ((c + a == 2) ? 7 : 8));
}

/* the two velocity components tangential to axis a, and the shear rate they
 * feed: a=0: (v2 -> s12, v3 -> s13); a=1: (v1 -> s12, v3 -> s23);
 * a=2: (v1 -> s13, v2 -> s23) */
__device__ __forceinline__ constexpr int tang(int a, int t)
{
  return a == 0 ? 1 + t : (a == 1 ? (t ? 2 : 0) : t);
}
__device__ __forceinline__ constexpr int shear_rate(int a, int t)
{
  return a == 0 ? (t ? 7 : 6) : (a == 1 ? (t ? 8 : 6) : (t ? 8 : 7));
}

/* upwind fluctuations for a face normal to axis A with sign s: from minus
 * and plus velocity v and stress column Sc (sigma[.][A]); izp = 1/(Zp- + Zp+),
 * izs = 1/(Zs- + Zs+) */
template <int A, class TC>
__device__ __forceinline__ void riemann_axis(TC s, const TC *vm, const TC *Sm,
                                             const TC *vp, const TC *Sp,
                                             TC Zpm, TC Zsm, TC Zpp, TC Zsp,
                                             TC izp, TC izs, TC *dv, TC *dT)
{
  /* traction T = sigma n = s * column A; n.T = s*s*sigma_AA */
  const TC dvn = (Zpp * (s * (vp[A] - vm[A])) + (Sp[A] - Sm[A])) * izp;
#pragma unroll
  for (int c = 0; c < 3; ++c)
  {
    if (c == A)
    {
      dv[c] = dvn * s;
      dT[c] = Zpm * dvn * s;
    }
    else
    {
      /* tangential: T_c = s sigma[c][A] */
      const TC dvt = (Zsp * (vp[c] - vm[c]) + s * (Sp[c] - Sm[c])) * izs;
      dv[c] = dvt;
      dT[c] = Zsm * dvt;
    }
  }
}

/* riemann_axis for an axis known only at run time */
template <class TC>
__device__ __forceinline__ void riemann_dispatch(int A, TC s, const TC *vm,
                                                 const TC *Sm, const TC *vp,
                                                 const TC *Sp, TC Zpm, TC Zsm,
                                                 TC Zpp, TC Zsp, TC izp, TC izs,
                                                 TC *dv, TC *dT)
{
  if (A == 0)
    riemann_axis<0, TC>(s, vm, Sm, vp, Sp, Zpm, Zsm, Zpp, Zsp, izp, izs, dv, dT);
  else if (A == 1)
    riemann_axis<1, TC>(s, vm, Sm, vp, Sp, Zpm, Zsm, Zpp, Zsp, izp, izs, dv, dT);
  else
    riemann_axis<2, TC>(s, vm, Sm, vp, Sp, Zpm, Zsm, Zpp, Zsp, izp, izs, dv, dT);
}

} // namespace bfamx

#endif
