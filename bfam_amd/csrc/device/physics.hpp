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

/* values a face node's lift adds to the rates of its volume node: the
 * velocities (0..2), the normal stresses (3..5) and the two shear stresses
 * that involve the face's axis (face_lift) */
constexpr int FACE_LIFT_N = 8;

/* FC: values per face node, 6 (the fluctuations dv, dT) or FACE_LIFT_N
 * (their lift, face_lift).  Where an element's nine fields are not a
 * multiple of 16 bytes (odd Np: N = 4, 6), an element block starts at any
 * 4-byte offset; its 16-byte LDS-DMA then starts at the 16-byte boundary
 * below it and ends at the one above, so the regions of the staged fields
 * and dq rows (QR, DR) hold up to 3 words before and after the QW values
 * and start on 16-byte boundaries. */
template <int N, int EPB, class T, bool DQ = false, int FC = FACE_LIFT_N> struct ElasticSmem
{
  static constexpr int n1 = N + 1, Np = n1 * n1 * n1, Nfp = n1 * n1;
  static constexpr int QW = EPB * 9 * Np;   /* staged fields */
  static constexpr bool SHIFT = (9 * Np) % 4 != 0;
  static constexpr int QR =
      SHIFT ? (QW * (int)sizeof(T) / 4 + 6) / 4 * 4 * 4 / (int)sizeof(T) : QW;
  static constexpr int DW = DQ ? QW : 0;    /* staged dq rows */
  static constexpr int DR = DQ ? QR : 0;
  static constexpr int FW = EPB * 6 * FC * Nfp; /* face values */
  /* LDS: fields, dq, face values, D, sin/cos tables */
  static constexpr size_t bytes = sizeof(T) * (QR + DR + FW + n1 * n1 + EPB * 6 * n1);
};

/* rate index of lift value 6 and 7 for a face of axis A */
__host__ __device__ constexpr int lift_shear0(int A) { return A == 2 ? 7 : 6; }
__host__ __device__ constexpr int lift_shear1(int A) { return A == 0 ? 7 : 8; }

/* The lift of one face node's fluctuations dv, dT, as the FACE_LIFT_N
 * values it adds to the rates (those elastic_lift adds; its other rate is
 * untouched):
 *   L[0..2] = lift dT / rho,
 *   L[3..5] = lift (lambda dvn + 2 mu s dv_c [c = axis]),
 *   L[6], L[7] = lift mu s dv_t for the shear stresses of the axis
 *                (rates lift_shear0(axis), lift_shear1(axis)). */
template <int AX, class T>
__device__ __forceinline__ void face_lift_axis(T s, T lift, T rhoi, T lam, T mu, const T *dv,
                                               const T *dT, T *L)
{
  L[0] = lift * dT[0] * rhoi;
  L[1] = lift * dT[1] * rhoi;
  L[2] = lift * dT[2] * rhoi;
  const T dvn = s * dv[AX];
  const T ld = lam * dvn;
  const T m2 = T(2) * mu * s;
  L[3] = lift * (ld + (AX == 0 ? m2 * dv[0] : T(0)));
  L[4] = lift * (ld + (AX == 1 ? m2 * dv[1] : T(0)));
  L[5] = lift * (ld + (AX == 2 ? m2 * dv[2] : T(0)));
  const T ms = mu * s;
  L[6] = lift * (ms * dv[AX == 0 ? 1 : 0]);
  L[7] = lift * (ms * dv[AX == 2 ? 1 : 2]);
}
/* face_lift_axis for an axis known only at run time (compile-time
 * indices in each branch: a run-time index into dv would put it in
 * scratch memory) */
template <class T>
__device__ __forceinline__ void face_lift(int axis, T s, T lift, T rhoi, T lam, T mu,
                                          const T *dv, const T *dT, T *L)
{
  if (axis == 0)
    face_lift_axis<0, T>(s, lift, rhoi, lam, mu, dv, dT, L);
  else if (axis == 1)
    face_lift_axis<1, T>(s, lift, rhoi, lam, mu, dv, dT, L);
  else
    face_lift_axis<2, T>(s, lift, rhoi, lam, mu, dv, dT, L);
}

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
