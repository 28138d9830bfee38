/* Physics helpers shared by the fused stage, unfused (plain) and reduction
 * kernels: precision traits, per-element table access, the elastic and
 * acoustic upwind fluctuations and lifts, manufactured-solution patterns. */
#pragma once

#include "launch.hpp"

namespace bfamx
{

template <class S, class R, class Cc, bool M = false> struct Prec
{
  using TS = S; /* state */
  using TR = R; /* rate register */
  using TC = Cc; /* arithmetic */
  /* derivative contractions in split-BF16x3 on the matrix cores */
  static constexpr bool MF = M;
};
using P64 = Prec<double, double, double>;
using P32R = Prec<double, float, float>;
using P32 = Prec<float, float, float>;
using P32M = Prec<float, float, float, true>;

/* per-precision copies of the element tables, as global pointers; SD is
 * SubDev, possibly qualified with the constant address space */
#define BFAMX_TAB(fn, dfield, ffield)                                          \
  template <class T, class SD>                                                 \
  __device__ __forceinline__ gptr<const T> fn(const SD &s)                     \
  {                                                                            \
    if constexpr (sizeof(T) == sizeof(double))                                 \
      return glob(s.dfield);                                                   \
    else                                                                       \
      return glob(s.ffield);                                                   \
  }
BFAMX_TAB(geo_tab, geo, geo_f)
BFAMX_TAB(mat_tab, mat, mat_f)
BFAMX_TAB(fmat_tab, fmat, fmat_f)
BFAMX_TAB(dr_tab, Dr, Dr_f)
BFAMX_TAB(src_tab, src_wt, src_wt_f)
BFAMX_TAB(mms_tab, mms_tab, mms_tab_f)
#undef BFAMX_TAB

/* which part of the stage a task body runs */
enum
{
  PART_FUSED = 0,   /* volume + faces + update (or RHS-only output) */
  PART_VOLUME = 1,  /* rhs = volume terms + sources (unfused path) */
  PART_SURFACE = 2  /* rhs += lifted face fluctuations (unfused path) */
};

template <int N, int EPB, class T, bool DQ = false> struct ElasticSmem
{
  static constexpr int n1 = N + 1, Np = n1 * n1 * n1, Nfp = n1 * n1;
  static constexpr int QW = EPB * 9 * Np;   /* staged fields */
  static constexpr int DW = DQ ? QW : 0;    /* staged dq rows */
  static constexpr int FW = EPB * 36 * Nfp; /* face fluctuations */
  /* LDS: fields, dq, face fluctuations, D, sin/cos tables */
  static constexpr size_t bytes = sizeof(T) * (QW + DW + FW + n1 * n1 + EPB * 6 * n1);
};

/* rate += lift * (dT/rho, lambda dvn I + mu (dv n^T + n dv^T)) */
template <class T>
__device__ __forceinline__ void elastic_lift(int axis, T s, T lift, T rhoi,
                                             T lam, T mu, const T *fl, T *rate)
{
  const T dv0 = fl[0], dv1 = fl[1], dv2 = fl[2];
  const T dvn = s * (axis == 0 ? dv0 : (axis == 1 ? dv1 : dv2));
  rate[0] += lift * fl[3] * rhoi;
  rate[1] += lift * fl[4] * rhoi;
  rate[2] += lift * fl[5] * rhoi;
  const T ld = lam * dvn;
  const T m2 = T(2) * mu * s;
  rate[3] += lift * (ld + (axis == 0 ? m2 * dv0 : T(0)));
  rate[4] += lift * (ld + (axis == 1 ? m2 * dv1 : T(0)));
  rate[5] += lift * (ld + (axis == 2 ? m2 * dv2 : T(0)));
  const T ms = mu * s;
  rate[6] += lift * (axis == 0 ? ms * dv1 : (axis == 1 ? ms * dv0 : T(0)));
  rate[7] += lift * (axis == 0 ? ms * dv2 : (axis == 2 ? ms * dv0 : T(0)));
  rate[8] += lift * (axis == 1 ? ms * dv2 : (axis == 2 ? ms * dv1 : T(0)));
}

template <class T>
__device__ __forceinline__ T mms_pattern(int f, const T *Sv, const T *Cv)
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


} // namespace bfamx
