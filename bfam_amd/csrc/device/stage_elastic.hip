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
 *   1. stream q[e] (NF*Np contiguous) from HBM and stage it in LDS: by
 *      LDS-DMA when LDS holds it exactly (FP32, FP64), through registers
 *      otherwise; the dq rows follow into LDS (FP32 rates) or registers
 *      (FP64), except on stages with A = 0, which do not read dq.  FP32 and
 *      FP64 store no dq: they read the previous stage's state instead and
 *      rebuild dq_{s-1} = (q_s - q_{s-1}) / B_{s-1}
 *   2. face tasks (one thread per face node): minus trace from LDS, plus trace
 *      (velocity and the stress column of the face's axis) gathered from the
 *      neighbour's q through nbr_elem/nbr_code (the on-device replacement of
 *      the glue trace exchange), upwind fluctuations (or, at the orders of
 *      NodeFaceLift, their lift: the eight rate increments of face_lift) ->
 *      LDS
 *   3. 18 directional derivatives per node from LDS (2(N+1) flop each),
 *      folded into the rates one axis at a time
 *   4. node owners lift their faces' fluctuations (or add the lift values)
 *   5. q_out (and dq, FP32 rates on FP64 state only) written: q ping-pongs
 *      so that neighbours read the stage input; nothing but the loop state
 *      is held across a barrier
 * The task, descriptor, clock and, for one-element blocks, the element
 * tables are read through the constant address space (scalar loads).
 */
#include "elastic_axis.hpp"
#include "mfma_deriv.hpp"
#include "physics.hpp"

namespace bfamx
{

/* Where the LGL lift of a face node is computed.  Face-side: the face task
 * turns its fluctuations into the FACE_LIFT_N rate increments (face_lift)
 * and the volume node adds those of the at most one face per axis it lies
 * on.  Node-side: the face task leaves its six fluctuations and the volume
 * node lifts each of its faces (elastic_lift); its wave then runs the lift
 * of every face any of its lanes lies on.  Face-side measured faster at
 * N = 4 and 7 (30^3 N=7 FP32: 0.398 -> 0.389 ms per stage, 48^3: 6.27 ->
 * 6.18 ms per step), slower at N = 3 and 6 (0.510 -> 0.529, 0.538 -> 0.572
 * ms), the same at N = 5. */
template <int N> struct NodeFaceLift
{
  static constexpr bool value = N == 4 || N == 5 || N == 7;
  static constexpr int FC = value ? FACE_LIFT_N : 6;
};

/* the dq rows of a task are staged in LDS by LDS-DMA (no registers held
 * while they load) for FP32 rates, where the larger LDS footprint still
 * lets as many blocks share a CU as the registers allow: three one-element
 * blocks (N = 6, 7), two of the multi-element blocks of N = 3..5, whose
 * register needs allow no more */
template <int N, int EPB, class P> struct NodeDq
{
  static constexpr bool value =
      sizeof(typename P::TR) == 4 &&
      (EPB == 1 ? 3 : 2) *
              ElasticSmem<N, EPB, typename P::TC, true, NodeFaceLift<N>::FC>::bytes <=
          160 * 1024;
};
template <int N, int EPB, class P>
using NodeSmem =
    ElasticSmem<N, EPB, typename P::TC, NodeDq<N, EPB, P>::value, NodeFaceLift<N>::FC>;
/* LDS of a block: the regions above, and for the matrix-core derivatives
 * the six derivative fields of one axis (FP32) */
template <int N, int EPB, class P> struct NodeLds
{
  static constexpr int MFW = P::MF ? 6 * EPB * (N + 1) * (N + 1) * (N + 1) : 0;
  static constexpr size_t bytes = NodeSmem<N, EPB, P>::bytes + sizeof(float) * MFW;
};

/*
 * Block-wide copy of `count` contiguous 32-bit words from global memory to
 * LDS with LDS-DMA (global_load_lds), W bytes per lane: the LDS image is the
 * global one.  Each wave's destination is uniform (lane l lands at
 * base + W*l); completion is tracked by vmcnt.  W = 16 needs count and both
 * addresses 16-byte aligned.  CPOL is the load's cache policy: 0, or
 * CPOL_NT for data no other block reads (streamed past L2 retention).
 */
constexpr int CPOL_NT = 2;

template <int W, int BLOCK, int CPOL = 0>
__device__ __forceinline__ void block_to_lds(gptr<const float> src, float *dst,
                                             int count)
{
  constexpr int per = W / 4;
  const int tid = threadIdx.x, wave = tid >> 6;
  for (int base = 0; base < count; base += BLOCK * per)
  {
    const int x = base + per * tid;
    auto *d = (__attribute__((address_space(3))) void *)(dst + base + per * 64 * wave);
    if (x < count)
    {
      if constexpr (W == 16)
        __builtin_amdgcn_global_load_lds(src + x, d, 16, 0, CPOL);
      else
        __builtin_amdgcn_global_load_lds(src + x, d, 4, 0, CPOL);
    }
  }
}

template <int N, int EPB, class P, int PART = PART_FUSED, class SD>
__device__ __forceinline__ void elastic_task(const SD &sd, const Task &tk,
                                             const StageCtx &cx, int mode,
                                             char *smem)
{
  using TS = typename P::TS;
  using TR = typename P::TR;
  using TC = typename P::TC;
  /* state re-read from LDS at the end when it is staged exactly */
  constexpr bool KEEP_Q = !__is_same(TS, TC);
  constexpr int n1 = N + 1, Np = n1 * n1 * n1, Nfp = n1 * n1, NF = 9;
  constexpr int BS = EPB * Np;
  constexpr int NT = EPB * 6 * Nfp;
  constexpr int TPT = (NT + BS - 1) / BS;
  using SM = NodeSmem<N, EPB, P>;
  /* state staged exactly (FP32, FP64): LDS-DMA straight into the staged
   * fields, 16 bytes per lane.  An element's words (QE state, DE dq) are
   * not a multiple of 4 at odd Np: the copy then starts at the 16-byte
   * boundary below the block, sh words before its first value, and the
   * staged fields begin sh words into their region (SM::SHIFT).  The
   * buffers are allocated 16 bytes longer than their values, so the
   * rounded-up end of the last element stays inside them (FP32 N = 4:
   * 1-element blocks 0.603 ms per 50^3 stage with 4-byte lanes) */
  constexpr bool DMA_Q = sizeof(TS) == sizeof(TC);
  constexpr bool DMA_DQ = SM::DW > 0;
  constexpr int QE = NF * Np * (int)sizeof(TS) / 4, DE = NF * Np;
  static_assert(!DMA_Q || QE % 4 == 0 || SM::SHIFT, "state rows need the shifted copy");
  static_assert(!DMA_DQ || DE % 4 == 0 || SM::SHIFT, "dq rows need the shifted copy");
  const int shq = DMA_Q && QE % 4 != 0 ? ((tk.e0 & 3) * (QE & 3)) & 3 : 0;
  const int shd = DMA_DQ && DE % 4 != 0 ? ((tk.e0 & 3) * (DE & 3)) & 3 : 0;

  const int tid = threadIdx.x;
  /* one element per block (N >= 6): the element index is uniform, and its
   * tables are read with scalar loads */
  using EP = typename ElemPtr<EPB == 1, TC>::type;
  const int le = EPB == 1 ? 0 : tid / Np, n = EPB == 1 ? tid : tid - (tid / Np) * Np;
  const bool act = tid < BS && le < tk.ne;
  const int e = tk.e0 + le;
  const int i = n % n1, j = (n / n1) % n1, k = n / (n1 * n1);

  TC *sq = reinterpret_cast<TC *>(smem + 4 * shq); /* [EPB][9][Np] */
  /* [EPB][9][Np] when DMA_DQ (FP32 rows) */
  TC *sdq = reinterpret_cast<TC *>(smem + sizeof(TC) * SM::QR + 4 * shd);
  /* face values: [EPB][6 faces][FC][Nfp] */
  constexpr bool FL = NodeFaceLift<N>::value;
  constexpr int FC = NodeFaceLift<N>::FC;
  TC *sfl = reinterpret_cast<TC *>(smem) + SM::QR + SM::DR;
  TC *sD = sfl + SM::FW;
  TC *smms = sD + n1 * n1;

  {
    gptr<const TC> Dg = dr_tab<TC>(sd);
    for (int t = tid; t < n1 * n1; t += blockDim.x)
      sD[t] = Dg[t];
  }
  if (PART != PART_SURFACE && sd.mms)
  {
    gptr<const TC> mt = mms_tab<TC>(sd) + (size_t)tk.e0 * 6 * n1;
    for (int t = tid; t < tk.ne * 6 * n1; t += blockDim.x)
      smms[t] = mt[t];
  }

  /* ---- 1. stream the element's state into LDS ---- */
  gptr<const TS> qin = glob(static_cast<const TS *>(sd.q[cx.parity]));
  TS qs[KEEP_Q ? NF : 1];
  if constexpr (DMA_Q)
  {
    block_to_lds<16, BS>((gptr<const float>)(qin + (size_t)tk.e0 * NF * Np) - shq,
                         reinterpret_cast<float *>(smem), (shq + tk.ne * QE + 3) & ~3);
  }
  else if (act)
  {
    gptr<const TS> qe = qin + (size_t)e * NF * Np + n;
#pragma unroll
    for (int f = 0; f < NF; ++f)
    {
      const TS v = qe[f * Np];
      if constexpr (KEEP_Q)
        qs[f] = v;
      sq[(le * NF + f) * Np + n] = (TC)v;
    }
  }
  /* the dq (or previous-state) rows load into LDS under the face and
   * volume work.  Only this block reads them, and the outputs are written
   * non-temporally: L2 then keeps the states that neighbouring blocks
   * gather their face traces from (FP32 48^3 N=7: 1.635 -> 1.548 ms per
   * stage) */
  const bool upd = !(mode == MODE_RHS || PART != PART_FUSED);
  /* a 2N-storage scheme starts every step with A = 0 (dq = dt L): the old
   * dq is not read then */
  const bool rd_dq = upd && cx.a != 0.0;
  /* FP32, FP64: the old dq is rebuilt from the state before this one,
   * which the other ping-pong buffer still holds (read here, at this
   * block's own nodes, before the same nodes are written below) */
  constexpr bool REC = sizeof(TS) == sizeof(TR);
  const bool rec = REC && cx.rec;
  const void *qprev = sd.q[cx.parity ^ 1];
  const void *dsrc = rec ? qprev : sd.dq;
  if constexpr (DMA_DQ)
  {
    if (rd_dq)
      block_to_lds<16, BS, CPOL_NT>(
          (gptr<const float>)glob(static_cast<const float *>(dsrc)) + (size_t)tk.e0 * DE - shd,
          reinterpret_cast<float *>(smem + sizeof(TC) * SM::QR), (shd + tk.ne * DE + 3) & ~3);
  }
  /* rates without room in LDS (FP64, and FP32 where a third block per CU
   * would not fit): this node's dq values load into registers instead, also
   * under the face and volume work */
  TR dqs[DMA_DQ ? 1 : NF];
  if constexpr (!DMA_DQ)
  {
    if (act && rd_dq)
    {
      gptr<const TR> dqe = glob(static_cast<const TR *>(dsrc)) + (size_t)e * NF * Np + n;
#pragma unroll
      for (int f = 0; f < NF; ++f)
        dqs[f] = __builtin_nontemporal_load(&dqe[f * Np]);
    }
  }
  /* LDS-DMA completes by vmcnt: every wave drains its own before the
   * barrier that publishes the staged state */
  if constexpr (DMA_Q || DMA_DQ)
    __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();

  /* ---- 2. faces: one task per face node, its lift values straight to LDS ---- */
  if constexpr (PART != PART_VOLUME)
  {
    gptr<const int32_t> nbe = glob(sd.nbr_elem);
    gptr<const uint8_t> nbc = glob(sd.nbr_code);
    gptr<const TC> mtab = mat_tab<TC>(sd);
    /* tasks axis-major (axis, element, side, face node), so that the
     * waves of a block mostly see faces of one axis: the flux and lift of
     * an axis run once per wave, not once per axis present in it */
    constexpr int TPA = EPB * 2 * Nfp;
#pragma unroll
    for (int r = 0; r < TPT; ++r)
    {
      const int t = tid + r * BS;
      const int A = t / TPA, ra = t - A * TPA;
      const int lt = EPB == 1 ? 0 : ra / (2 * Nfp);
      if (tid < BS && t < NT && lt < tk.ne)
      {
        const int rs = ra - lt * 2 * Nfp;
        const int side = rs / Nfp, fn = rs - side * Nfp;
        const int f = 2 * A + side;
        const int et = tk.e0 + lt;
        const int a = fn % n1, b = fn / n1;
        const int nm = face_node<3>(f, n1, a, b);
        const int32_t ep = nbe[(size_t)et * 6 + f];
        const int code = nbc[(size_t)et * 6 + f];
        TC *o = sfl + (lt * 6 + f) * FC * Nfp + fn;
        EP Mm = (EP)mtab + (size_t)et * MAT_STRIDE;
        const TC lf = (EP(geo_tab<TC>(sd)) + (size_t)et * GEO_STRIDE)[GEO_LX + A];
        TC dv[3], dT[3];
        if (code & NBR_GLUED)
        {
          /* fluctuations already on this face node (mortar kernel) */
          gptr<const TC> gb = glob(static_cast<const TC *>(sd.glue_buf)) + ep + fn;
#pragma unroll
          for (int c = 0; c < 3; ++c)
          {
            dv[c] = gb[c * Nfp];
            dT[c] = gb[(3 + c) * Nfp];
          }
        }
        else
        {
          gptr<const TS> qp = qin + (size_t)ep * NF * Np + nbr_node(code, f, n1, nm);
          const TC *u = sq + lt * NF * Np + nm;
          TC sv, ss;
          nbr_mirror<TC>(code, &sv, &ss);
          /* only the velocity and the stress column of the face's axis enter
           * the flux: six of the nine neighbour values are gathered */
          TC vm[3], Sm[3], vp[3], Sp[3];
#pragma unroll
          for (int c = 0; c < 3; ++c)
          {
            const int col = sfield(c, A);
            vm[c] = u[c * Np];
            Sm[c] = u[col * Np];
            vp[c] = sv * (TC)qp[c * Np];
            Sp[c] = ss * (TC)qp[col * Np];
          }
          gptr<const TC> F = fmat_tab<TC>(sd) + ((size_t)et * 6 + f) * FMAT_STRIDE;
          riemann_dispatch<TC>(A, (f & 1) ? TC(1) : TC(-1), vm, Sm, vp, Sp, Mm[MAT_ZP],
                               Mm[MAT_ZS], F[FMAT_ZP], F[FMAT_ZS], F[FMAT_IZP],
                               F[FMAT_IZS], dv, dT);
        }
        if constexpr (FL)
        {
          TC L[FACE_LIFT_N];
          face_lift<TC>(A, (f & 1) ? TC(1) : TC(-1), lf, Mm[MAT_RHOI], Mm[MAT_LAM],
                        Mm[MAT_MU], dv, dT, L);
#pragma unroll
          for (int c = 0; c < FACE_LIFT_N; ++c)
            o[c * Nfp] = L[c];
        }
        else
        {
#pragma unroll
          for (int c = 0; c < 3; ++c)
          {
            o[c * Nfp] = dv[c];
            o[(3 + c) * Nfp] = dT[c];
          }
        }
      }
    }
    __syncthreads();
  }

  TC rate[NF];
  /* matrix-core derivatives: every thread of the block takes part (the
   * tiles go to the full wavefronts), one axis at a time through a six-field
   * LDS buffer; each node then folds its six values of the axis into the
   * rates exactly as the contraction below does */
  if constexpr (P::MF && PART == PART_FUSED)
  {
    float *der = reinterpret_cast<float *>(smem + NodeSmem<N, EPB, P>::bytes);
    EP M0 = (EP)mat_tab<TC>(sd) + (size_t)(act ? e : tk.e0) * MAT_STRIDE;
    EP G0 = (EP)geo_tab<TC>(sd) + (size_t)(act ? e : tk.e0) * GEO_STRIDE;
    const TC lam = M0[MAT_LAM], mu = M0[MAT_MU], l2m = lam + TC(2) * mu;
    const float *dr = der + le * Np + n;
    constexpr int SW = EPB * Np;
    {
      constexpr int fx[6] = {0, 1, 2, 3, 6, 7};
      mf_axis<N, EPB, 0>(sq, sD, der, fx);
      __syncthreads();
      const TC g = G0[GEO_GX];
      const TC v1 = g * dr[0], v2 = g * dr[SW], v3 = g * dr[2 * SW], s11 = g * dr[3 * SW],
               s12 = g * dr[4 * SW], s13 = g * dr[5 * SW];
      rate[0] = s11;
      rate[1] = s12;
      rate[2] = s13;
      rate[3] = l2m * v1;
      rate[4] = lam * v1;
      rate[5] = lam * v1;
      rate[6] = mu * v2;
      rate[7] = mu * v3;
      rate[8] = 0;
      __syncthreads();
    }
    {
      constexpr int fy[6] = {0, 1, 2, 4, 6, 8};
      mf_axis<N, EPB, 1>(sq, sD, der, fy);
      __syncthreads();
      const TC g = G0[GEO_GY];
      const TC v1 = g * dr[0], v2 = g * dr[SW], v3 = g * dr[2 * SW], s22 = g * dr[3 * SW],
               s12 = g * dr[4 * SW], s23 = g * dr[5 * SW];
      rate[0] += s12;
      rate[1] += s22;
      rate[2] += s23;
      rate[3] += lam * v2;
      rate[4] += l2m * v2;
      rate[5] += lam * v2;
      rate[6] += mu * v1;
      rate[8] += mu * v3;
      __syncthreads();
    }
    {
      constexpr int fz[6] = {0, 1, 2, 5, 7, 8};
      mf_axis<N, EPB, 2>(sq, sD, der, fz);
      __syncthreads();
      const TC g = G0[GEO_GZ];
      const TC v1 = g * dr[0], v2 = g * dr[SW], v3 = g * dr[2 * SW], s33 = g * dr[3 * SW],
               s13 = g * dr[4 * SW], s23 = g * dr[5 * SW];
      rate[0] += s13;
      rate[1] += s23;
      rate[2] += s33;
      rate[3] += lam * v3;
      rate[4] += lam * v3;
      rate[5] += l2m * v3;
      rate[7] += mu * v1;
      rate[8] += mu * v2;
    }
  }
  if (!act)
    return;

  EP M = (EP)mat_tab<TC>(sd) + (size_t)e * MAT_STRIDE;
  const TC rhoi = M[MAT_RHOI], lam = M[MAT_LAM], mu = M[MAT_MU];
  EP G = (EP)geo_tab<TC>(sd) + (size_t)e * GEO_STRIDE;
  if (PART == PART_SURFACE)
  {
    gptr<const double> rh = glob((const double *)sd.rhs) + (size_t)e * NF * Np + n;
#pragma unroll
    for (int f = 0; f < NF; ++f)
      rate[f] = (TC)rh[f * Np];
  }
  else if constexpr (P::MF)
  {
    rate[0] *= rhoi;
    rate[1] *= rhoi;
    rate[2] *= rhoi;
  }
  else
  {
    /* ---- 3. volume: 18 directional derivatives, folded into the rates one
     * axis at a time so only six accumulators are live ---- */
    const TC *u = sq + le * NF * Np;
    const TC l2m = lam + TC(2) * mu;
    {
      TC v1 = 0, v2 = 0, v3 = 0, s11 = 0, s12 = 0, s13 = 0;
#pragma unroll 2
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
#pragma unroll 2
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
#pragma unroll 2
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

  /* ---- 4. lift: face-side, the lift values of the node's faces (at
   * most one per axis); node-side, the node lifts its faces' fluctuations
   * in the fixed order -x, +x, -y, +y, -z, +z ---- */
  if constexpr (PART != PART_VOLUME && FL)
  {
    const TC *F = sfl + le * 6 * FC * Nfp;
#define BFAMX_LIFT(A, c_, fn_)                                                 \
  if ((c_) == 0 || (c_) == N)                                                  \
  {                                                                            \
    const TC *Ff = F + (2 * (A) + ((c_) == N)) * FC * Nfp + (fn_);              \
    _Pragma("unroll") for (int c = 0; c < 6; ++c) rate[c] += Ff[c * Nfp];      \
    rate[lift_shear0(A)] += Ff[6 * Nfp];                                       \
    rate[lift_shear1(A)] += Ff[7 * Nfp];                                       \
  }
    BFAMX_LIFT(0, i, j + n1 * k)
    BFAMX_LIFT(1, j, i + n1 * k)
    BFAMX_LIFT(2, k, i + n1 * j)
#undef BFAMX_LIFT
  }
  else if constexpr (PART != PART_VOLUME)
  {
    const TC *F = sfl + le * 6 * FC * Nfp;
    TC flv[6];
#define BFAMX_LIFT(face, cond, fn_)                                            \
  if (cond)                                                                    \
  {                                                                            \
    _Pragma("unroll") for (int c = 0; c < 6; ++c) flv[c] =                     \
        F[((face)*FC + c) * Nfp + (fn_)];                                      \
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
    const TC *T = smms + le * 6 * n1;
    const TC Sv[3] = {T[0 * n1 + i], T[2 * n1 + j], T[4 * n1 + k]};
    const TC Cv[3] = {T[1 * n1 + i], T[3 * n1 + j], T[5 * n1 + k]};
    const TC sw = (TC)cx.sw, cw = (TC)cx.cw;
#pragma unroll
    for (int f = 0; f < NF; ++f)
      rate[f] += (TC)sd.mms_F[f] * mms_pattern<TC>(f, Sv, Cv) * (f < 3 ? sw : cw);
  }

  /* ---- 5. output ---- */
  if (!upd)
  {
    gptr<double> rh = glob(sd.rhs) + (size_t)e * NF * Np + n;
#pragma unroll
    for (int f = 0; f < NF; ++f)
      rh[f * Np] = (double)rate[f];
    return;
  }
  gptr<TS> qo = glob(static_cast<TS *>(sd.q[cx.parity ^ 1])) + (size_t)e * NF * Np + n;
  gptr<TR> dqe = glob(static_cast<TR *>(sd.dq)) + (size_t)e * NF * Np + n;
  const TR ca = (TR)cx.a, cdt = (TR)cx.dt;
  const TS cb = (TS)cx.b;
#pragma unroll
  for (int f = 0; f < NF; ++f)
  {
    TS qv;
    if constexpr (KEEP_Q)
      qv = qs[f];
    else
      qv = (TS)sq[(le * NF + f) * Np + n];
    TR dqo = 0;
    if (rd_dq)
    {
      if constexpr (DMA_DQ)
        dqo = (TR)sdq[(le * NF + f) * Np + n];
      else
        dqo = dqs[f];
      /* rebuilt: B_{s-1} dq_{s-1} = q_s - q_{s-1}; a holds A_s / B_{s-1} */
      if constexpr (REC)
        if (rec)
          dqo = (TR)qv - dqo;
    }
    const TR d = ca * dqo + cdt * (TR)rate[f];
    /* stored only where the next stage reads it and cannot rebuild it */
    if (!cx.dq_dead && !rec)
      __builtin_nontemporal_store(d, &dqe[f * Np]);
    __builtin_nontemporal_store((TS)(qv + cb * (TS)d), &qo[f * Np]);
  }
}

/* waves per SIMD the register allocation must leave room for.  For the
 * one-element blocks (FP32 at N = 3..7, FP64 at N = 4..7, FP32 rates at
 * N = 5..7): 6 (80 registers) keeps three 512-thread blocks per CU for the
 * FP32 stage at N = 7 and runs the FP32 one-wavefront blocks of N = 3
 * (0.441 against 0.447 ms per 60^3 stage at 96 registers); 5 (96
 * registers, no spill) suits FP32 at N = 4, 5 (N = 5: 0.459 against 0.479
 * at 80 registers); FP32_RHS and FP64 take 4 (forcing FP32_RHS to three
 * N = 7 blocks spilled three registers and measured slower, 3.01 against
 * 2.88 ms per 48^3 stage).  Multi-element blocks are left to the
 * compiler: capping the four-element FP32 blocks of N = 4 for a third block
 * per CU measured slower (0.594 -> 0.659), as did 8 for FP32 at N = 6 (five
 * blocks per CU, 0.432 -> 0.539). */
template <int N, int EPB, class P> struct NodeWaves
{
  static constexpr int value =
      P::MF ? 4
            : (EPB == 1 ? (((N == 4 || N == 5) && __is_same(P, P32)) ? 5
                           : (sizeof(typename P::TS) == 4 ? 6 : 4))
                        : 1);
};

/* elements per block of the fused node-parallel kernel.  A node task holds
 * up to 512 / (N+1)^3 elements (elastic_epb); a block of fewer elements
 * takes a task in passes.  Smaller blocks let more of them share a CU
 * (the LDS of a block scales with its elements), ms per stage on bricks of
 * about 124 M DOF, one-block-per-task grid:
 *   N = 5, one element (216 threads): FP32 0.541 -> 0.483, FP32 rates
 *   0.852 -> 0.739, FP64 1.009 -> 0.795;
 *   N = 3: FP32 one element (one wavefront) 0.482 -> 0.444, FP64 two
 *   elements 1.05 -> 0.806 (its line kernel: 1.017);
 *   N = 4: FP64 one element 1.28 -> 0.948 (its line kernel: 1.188); FP32
 *   one element at 96 registers 0.590 -> 0.572 (resident grid; at 80
 *   registers 0.607, at 114 0.610; two elements 0.626).
 * FP32 rates at N = 3, 4 with one element measured slower taking a task
 * in passes (0.708 -> 0.759, 0.894 -> 1.007) and faster with one block per
 * element (NodeSplit): 0.705 -> 0.642, 0.868 -> 0.813. */
template <int N, class P> struct NodeEpb
{
  static constexpr bool f32 = __is_same(P, P32) || __is_same(P, P32R), f64 = __is_same(P, P64);
  static constexpr int value =
      N == 3 ? (f32 ? 1 : (f64 ? 2 : 8)) : (N == 4 ? ((f64 || f32) ? 1 : 4) : 1);
};
/* the other block size the autotuner tries at N = 3..5 (kernel kinds 3
 * and 4): the task-sized blocks where NodeEpb is smaller, one element
 * where it is the task size, two elements at N = 5 */
template <int N, class P> struct NodeEpbAlt
{
  static constexpr int TE = 512 / ((N + 1) * (N + 1) * (N + 1)) < 1 ? 1 : 512 / ((N + 1) * (N + 1) * (N + 1));
  static constexpr int value =
      N == 5 ? 2 : (N > 5 ? 1 : (NodeEpb<N, P>::value < TE ? TE : 1));
};

/* blocks per node task on a one-block-per-task grid: the task's elements
 * over EPB (each block taking EPB of them) where that measured faster than
 * one block taking the whole task in passes, 1 elsewhere.  FP64 at N = 3
 * (two-element blocks): 0.805 -> 0.789 ms per stage.  Each XCD's
 * blocks then cover 2x2x2 elements of a Morton group instead of 4x4x4
 * (N = 3), and fewer elements are in flight per XCD.  ms per stage on
 * bricks of about 124 M DOF: FP32 N = 3 0.443 -> 0.412-0.414, N = 4
 * 0.548-0.552 -> 0.523-0.530.  FP32 at N = 5 measured slower with
 * one-element tasks (0.455 -> 0.464), as did the resident grids; the FP64
 * N = 4 form spilled 12 bytes per lane and took 0.866 -> 0.946 ms.  An
 * order launched beside a larger one of its batch keeps whole tasks
 * (LaunchArgs::whole_tasks), as the autotuner then times it too: split
 * there, its many small blocks crowded the larger order's (adapted mesh,
 * orders 7/5/3: 0.555-0.567 ms per step, 0.550-0.553 with whole tasks). */
template <int N, int EPB, class P> struct NodeSplit
{
  static constexpr int TE = 512 / ((N + 1) * (N + 1) * (N + 1)) < 1 ? 1 : 512 / ((N + 1) * (N + 1) * (N + 1));
  static constexpr int value =
      (EPB == 1 && (__is_same(P, P32) || __is_same(P, P32R)) && (N == 3 || N == 4)) ||
              (EPB == 2 && __is_same(P, P64) && N == 3)
          ? TE / EPB
          : 1;
};

/* SPLIT: the launch of a NodeSplit order with NS blocks per task, block v
 * taking elements EPB (v % NS) .. EPB (v % NS) + EPB - 1 of task v / NS */
template <int N, int EPB, class P, bool SPLIT = false>
__global__ __launch_bounds__(EPB *(N + 1) * (N + 1) * (N + 1))
__attribute__((amdgpu_waves_per_eu(NodeWaves<N, EPB, P>::value))) void
elastic_stage_kernel(LaunchArgs a)
{
  __shared__ __align__(16) char smem[NodeLds<N, EPB, P>::bytes];
  /* elements of a node task */
  constexpr int TE = 512 / ((N + 1) * (N + 1) * (N + 1)) < 1 ? 1 : 512 / ((N + 1) * (N + 1) * (N + 1));
  static_assert(TE % EPB == 0, "a block takes whole passes of a task");
  int stride, end;
  if constexpr (SPLIT)
  {
    constexpr int NS = NodeSplit<N, EPB, P>::value;
    static_assert(NS > 1 && NS * EPB == TE, "split launches take EPB elements per block");
    for (int vi = first_task(a, a.ntasks * NS, &stride, &end); vi < end; vi += stride)
    {
      const int ti = vi / NS, s0 = (vi - ti * NS) * EPB;
      Task tk = load_task(a.tasks, ti);
      if (s0 < tk.ne)
      {
        const auto &sd = *(cglob(a.subs) + tk.sub);
        const StageCtx cx = stage_ctx(sd, a);
        tk.e0 += s0;
        tk.ne = tk.ne - s0 < EPB ? tk.ne - s0 : EPB;
        elastic_task<N, EPB, P>(sd, tk, cx, a.mode, smem);
        __syncthreads();
      }
    }
    return;
  }
  for (int ti = first_task(a, &stride, &end); ti < end; ti += stride)
  {
    const Task tk = load_task(a.tasks, ti);
    const auto &sd = *(cglob(a.subs) + tk.sub);
    const StageCtx cx = stage_ctx(sd, a);
    if constexpr (EPB == TE)
    {
      elastic_task<N, EPB, P>(sd, tk, cx, a.mode, smem);
      __syncthreads();
    }
    else
      for (int s0 = 0; s0 < tk.ne; s0 += EPB)
      {
        Task pt = tk;
        pt.e0 = tk.e0 + s0;
        pt.ne = tk.ne - s0 < EPB ? tk.ne - s0 : EPB;
        elastic_task<N, EPB, P>(sd, pt, cx, a.mode, smem);
        __syncthreads();
      }
  }
}

/* the launch of order N with EPB elements per block, or the occupancy
 * calculator's resident blocks per CU for it */
template <int N, int EPB, class P>
static hipError_t launch_ne(const LaunchArgs &a, int grid, hipStream_t s, unsigned dyn)
{
  constexpr int NS = NodeSplit<N, EPB, P>::value;
  /* one block per task (not a resident grid): one block per element */
  if constexpr (NS > 1)
    if (grid >= a.ntasks && !a.whole_tasks)
    {
      const int vt = a.ntasks * NS;
      grid = a.xcd_remap == 1 && a.nxcd > 1 ? a.nxcd * ((vt + a.nxcd - 1) / a.nxcd) : vt;
      hipLaunchKernelGGL((elastic_stage_kernel<N, EPB, P, true>), dim3(grid),
                         dim3(EPB * (N + 1) * (N + 1) * (N + 1)), dyn, s, a);
      return hipGetLastError();
    }
  hipLaunchKernelGGL((elastic_stage_kernel<N, EPB, P>), dim3(grid),
                     dim3(EPB * (N + 1) * (N + 1) * (N + 1)), dyn, s, a);
  return hipGetLastError();
}
template <int N, int EPB, class P>
static hipError_t occupancy_ne(int *blocks, unsigned dyn)
{
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(
      blocks, elastic_stage_kernel<N, EPB, P>, EPB * (N + 1) * (N + 1) * (N + 1), dyn);
}

/* alt: the block size of NodeEpbAlt instead of NodeEpb */
template <class P>
static hipError_t launch_n(int N, int alt, const LaunchArgs &a, int grid, hipStream_t s,
                           unsigned dyn)
{
  switch (N)
  {
#define BFAMX_CASE(NN)                                                         \
  case NN:                                                                     \
    return alt ? launch_ne<NN, NodeEpbAlt<NN, P>::value, P>(a, grid, s, dyn)    \
               : launch_ne<NN, NodeEpb<NN, P>::value, P>(a, grid, s, dyn);
    BFAMX_CASE(3)
    BFAMX_CASE(4)
    BFAMX_CASE(5)
    BFAMX_CASE(6)
    BFAMX_CASE(7)
#undef BFAMX_CASE
  default: return hipErrorInvalidValue;
  }
}

template <class P>
static hipError_t occupancy_node(int N, int alt, int *blocks, unsigned dyn)
{
  switch (N)
  {
#define BFAMX_CASE(NN)                                                         \
  case NN:                                                                     \
    return alt ? occupancy_ne<NN, NodeEpbAlt<NN, P>::value, P>(blocks, dyn)     \
               : occupancy_ne<NN, NodeEpb<NN, P>::value, P>(blocks, dyn);
    BFAMX_CASE(3)
    BFAMX_CASE(4)
    BFAMX_CASE(5)
    BFAMX_CASE(6)
    BFAMX_CASE(7)
#undef BFAMX_CASE
  default: return hipErrorInvalidValue;
  }
}

hipError_t node_blocks_per_cu(int prec, int N, int *blocks, unsigned dyn, int alt)
{
  switch (prec)
  {
  case PREC_FP32: return occupancy_node<P32>(N, alt, blocks, dyn);
  case PREC_FP32_RHS: return occupancy_node<P32R>(N, alt, blocks, dyn);
  case PREC_FP32_BF16X3: return occupancy_node<P32M>(N, alt, blocks, dyn);
  default: return occupancy_node<P64>(N, alt, blocks, dyn);
  }
}

hipError_t launch_elastic_stage(int prec, int N, const LaunchArgs &a, int grid,
                                hipStream_t s, unsigned dyn, int alt)
{
  switch (prec)
  {
  case PREC_FP32: return launch_n<P32>(N, alt, a, grid, s, dyn);
  case PREC_FP32_RHS: return launch_n<P32R>(N, alt, a, grid, s, dyn);
  case PREC_FP32_BF16X3: return launch_n<P32M>(N, alt, a, grid, s, dyn);
  default: return launch_n<P64>(N, alt, a, grid, s, dyn);
  }
}

/* unfused FP64 form: the same element bodies run as a volume kernel and a
 * surface kernel, each a separate pass over HBM through the rhs array */
template <int N, int EPB, int PART>
__global__ __launch_bounds__(EPB *(N + 1) * (N + 1) * (N + 1)) void
elastic_part_kernel(LaunchArgs a)
{
  __shared__ __align__(16) char smem[NodeLds<N, EPB, P64>::bytes];
  int stride, end;
  for (int ti = first_task(a, &stride, &end); ti < end; ti += stride)
  {
    const Task tk = load_task(a.tasks, ti);
    const auto &sd = *(cglob(a.subs) + tk.sub);
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
