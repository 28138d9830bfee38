/*
 * Fused LSRK stage for 3-D isotropic velocity-stress elasticity, "line"
 * formulation: one wavefront per element (N = 7; 64 / (N+1)^2 elements per
 * wave for smaller N), each lane owning one line of N+1 nodes.
 *
 * Same method and result as the node-parallel body in stage_elastic.hip
 * (strong-form LGL DG, exact-Riemann upwind fluctuations, LGL lift, LSRK
 * 2N-storage update; see there for the references), organised for CDNA
 * wavefronts:
 *   - the state moves between HBM and LDS as coalesced rows (lane l takes
 *     nodes VW*(l + 64 r) .. +VW-1 of each field row, 16-byte accesses);
 *     LDS holds it densely in the arithmetic type (node (i,j,k) at
 *     i + n1 j + n1^2 k, swizzled at N = 7 so that every pass is free of
 *     bank conflicts, see LineCfg);
 *   - lane l = (a, b) = (l % n1, l / n1) of its element's n1 x n1 lines;
 *   - x pass: lane owns the i-line (j, k) = (a, b) and accumulates the nine
 *     rates of that line in registers; the contraction with D uses the
 *     wave-uniform D held in scalar registers;
 *   - y pass: lane owns the j-line (i, k) = (a, b);
 *   - z pass: lane owns the k-line (i, j) = (a, b);
 *   - in each pass the lane also owns face node (a, b) of the two faces normal
 *     to that axis, which are the two ends of its line, so the gather of the
 *     neighbour trace, the Riemann solve and the lift are spread evenly over
 *     all lanes with no divergence;
 *   - y and z contributions (six per axis: the three traction components,
 *     the normal velocity and the two tangential velocities, each
 *     already including the face lift) return to the x-line owners through
 *     a one-field LDS transpose, where the material combines them into the
 *     nine rates;
 *   - the rates go back to row order through the same transpose buffer, one
 *     field at a time, and q is updated with coalesced row accesses: the
 *     old dq, or for FP32 and FP64 the previous state it is rebuilt from,
 *     is fetched ahead of use; only FP32_RHS stores dq.
 * Nothing crosses a workgroup barrier: each wave only touches its own LDS
 * region, and LDS operations of one wave execute in issue order.
 */
#include <type_traits>

#include "elastic_axis.hpp"
#include "physics.hpp"

namespace bfamx
{

template <int N, class TL, class TC> struct LineCfg
{
  static constexpr int n1 = N + 1;
  static constexpr int NL = n1 * n1;  /* lines per element */
  static constexpr int Np = n1 * NL;
  static constexpr int EPW = 64 / NL; /* elements per wavefront */
  /* k stride of the LDS layout.  Padding it to n1^2 + n1 spreads the
   * strided j-line reads over more banks, but costs 1/n1 more LDS per wave:
   * at N = 7 the dense layout fits 7 instead of 6 waves per CU and measured
   * 2.5 % faster (2.30 against 2.36 ms per 48^3 stage), so it is dense */
  static constexpr int S = NL;
  /* At N = 7 the words of a field are swizzled: node (i, j, k) sits at
   * 64 k + 8 (j ^ k) + (i ^ 4 (k & 1)).  Each k-plane keeps its 64 words and
   * each half x-line (i = 0..3 or 4..7) its 4 consecutive words, so rows
   * still move as 16-byte vectors, but the lanes of every pass -- (j, k) for
   * x-lines, (i, k) for y-lines, (i, j) for z-lines -- now fall on distinct
   * banks (the dense layout puts the y-line reads of a wave on 8 banks) */
  static constexpr bool swz = (n1 == 8);
  static constexpr int FS = S * n1;    /* words per field of one element */
  static constexpr int FST = EPW * FS; /* field stride: the wave's elements */
  /* per wavefront: 9 staged state fields (type TL), field-major over the
   * wave's elements so that a coalesced row lands at consecutive words.
   * The transposes of the y and z passes and of the output reuse the
   * region of sigma_11, which only the x pass reads (its row values are
   * kept in registers for the update), so no separate buffer is needed */
  static constexpr size_t bytes = EPW * FS * 9 * sizeof(TL);
};

/* vector width of the coalesced row accesses: 16-byte loads when every
 * element block of NF*Np values keeps that alignment */
template <int N, class TS> struct CoalVW
{
  static constexpr int np = (N + 1) * (N + 1) * (N + 1);
  static constexpr int w = 16 / (int)sizeof(TS);
  static constexpr int value = (np % w == 0) ? w : 1;
};

/* the next task of a persistent wavefront, whose rows it starts loading
 * while the current task writes its output */
template <class TS> struct NextRows
{
  gptr<const TS> q; /* state buffer the next task reads */
  int e0, ne;
  bool valid;
};

/*
 * Field f of a task's state straight from HBM into LDS (global_load_lds, no
 * register staging): lane l moves nodes VW*(l + 64 r) .. +VW-1 of the field
 * row group, counted flat over the task's elements, to the same word offsets
 * of the field's LDS region -- the LDS-DMA unit writes lane l's VW words at
 * base + VW*l, which is exactly that offset.  Completion is tracked by
 * vmcnt; readers wait for it.
 */
template <int N, class TS>
__device__ __forceinline__ void rows_to_lds(gptr<const TS> q, int e0, int ne, int f,
                                            TS *field_lds)
{
  constexpr int n1 = N + 1, Np = n1 * n1 * n1, EPW = 64 / (n1 * n1);
  constexpr int VW = CoalVW<N, TS>::value;
  constexpr int ROWS = (EPW * Np + 64 * VW - 1) / (64 * VW);
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int r = 0; r < ROWS; ++r)
  {
    const int x = VW * (lane + 64 * r);
    if (x < ne * Np)
    {
      const int lel = x / Np, n = x - lel * Np;
      /* the node whose words land at LDS word n (inverse of the swizzle) */
      int sn = n;
      if constexpr (LineCfg<N, TS, TS>::swz)
      {
        const int k = n >> 6;
        sn = 64 * k + 8 * (((n >> 3) & 7) ^ k) + ((n & 4) ^ ((k & 1) << 2));
      }
      gptr<const TS> src = q + ((size_t)(e0 + lel) * 9 + f) * Np + sn;
      if constexpr (VW * sizeof(TS) == 16)
        __builtin_amdgcn_global_load_lds(
            src, (__attribute__((address_space(3))) void *)(field_lds + 64 * VW * r), 16,
            0, 0);
      else if constexpr (VW * sizeof(TS) == 4)
        __builtin_amdgcn_global_load_lds(
            src, (__attribute__((address_space(3))) void *)(field_lds + 64 * VW * r), 4,
            0, 0);
    }
  }
}

/*
 * The six per-axis raw quantities of a line, with face lifts folded in:
 *   Q[0..2] = g D sigma[.][A] (+ lift dT)        -> rates 0..2 (times 1/rho)
 *   Q[3]    = g D v_A        (+ lift s dv_A)     -> normal-stress rates
 *   Q[4,5]  = g D v_tang     (+ lift s dv_tang)  -> shear rates
 * w[f][m]: the line's values of the six fields (order of qfield()).
 */
template <int A>
__device__ __forceinline__ constexpr int qfield(int c)
{
  return c < 3 ? sfield(c, A) : (c == 3 ? A : tang(A, c - 4));
}

/* fold the rate contribution of raw quantity c of axis A at one node;
 * coefficients already carry the 2/h_A factor */
template <int A, class TC>
__device__ __forceinline__ void fold(int c, TC q, TC g, TC glam, TC gl2m,
                                     TC gmu, TC *rate /* 9 */)
{
  if (c < 3)
    rate[c] += g * q;
  else if (c == 3)
  {
    rate[3] += (A == 0 ? gl2m : glam) * q;
    rate[4] += (A == 1 ? gl2m : glam) * q;
    rate[5] += (A == 2 ? gl2m : glam) * q;
  }
  else
    rate[shear_rate(A, c - 4)] += gmu * q;
}

/*
 * One task.  When the state type is the LDS type (FP32, FP64) its rows reach
 * LDS by LDS-DMA; rows_ready says a previous task of this wavefront already
 * issued them, and a valid nx gets its rows issued field by field as this
 * task's output frees each field's LDS region (persistent wavefronts).
 */
template <int N, class P, int PART = PART_FUSED, class SD>
__device__ __forceinline__ void elastic_line_task(const SD &sd,
                                                  const Task &tk,
                                                  const StageCtx &cx, int mode,
                                                  char *lds, bool rows_ready,
                                                  const NextRows<typename P::TS> &nx)
{
  using TS = typename P::TS;
  using TR = typename P::TR;
  using TC = typename P::TC;
  /* LDS copy of the state in the arithmetic type; the update takes q_old
   * from that copy when it is exact (TS == TC), else from global memory */
  using TL = TC;
  constexpr bool q_from_lds = std::is_same<TS, TL>::value;
  using CF = LineCfg<N, TL, TC>;
  constexpr int n1 = CF::n1, NL = CF::NL, Np = CF::Np, S = CF::S, FS = CF::FS;
  constexpr int FST = CF::FST;
  /* LDS-DMA moves 4 or 16 bytes per lane.  It is used for the FP32 state;
   * the FP64 stage measured 10 % slower with it than with staging through
   * registers (5.77 against 6.34 ms per 48^3 N=7 stage), and FP64 rows of
   * odd-sized elements (one double per lane) cannot use it */
  constexpr int dma_bytes = CoalVW<N, TS>::value * (int)sizeof(TS);
  constexpr bool dma = std::is_same<TS, TL>::value && sizeof(TS) == 4 &&
                       (dma_bytes == 4 || dma_bytes == 16);
  constexpr int NF = 9;

  const int lane = threadIdx.x & 63;
  const int le = lane / NL, L = lane - (lane / NL) * NL;
  const bool act = le < tk.ne;
  const int e = tk.e0 + (act ? le : 0);
  const int pa = L % n1, pb = L / n1; /* (a, b) of this lane in every pass */
  TL *sq = reinterpret_cast<TL *>(lds) + le * FS; /* staged state, field stride FST */
  static_assert(std::is_same<TL, TC>::value, "the transposes share the LDS state");
  constexpr int TBF = 3; /* sigma_11: its region is the transpose buffer after x */
  TC *tb = reinterpret_cast<TC *>(reinterpret_cast<TL *>(lds) + TBF * FST) +
           le * FS; /* one-field transpose buffer */

  gptr<const TS> qin = glob(static_cast<const TS *>(sd.q[cx.parity]));
  /* element tables and D: through the constant address space in FP32,
   * where one element per wave (N = 7) makes them scalar loads.  FP64 reads
   * them with vector loads: D alone would take 128 scalar registers, and the
   * scalar spills measured 20 % slower (6.5 against 5.4 ms per 48^3 stage) */
  using EP = typename ElemPtr<sizeof(TC) == 4, TC>::type;
  EP mtab = (EP)mat_tab<TC>(sd);
  EP fmt = (EP)fmat_tab<TC>(sd) + (size_t)e * 6 * FMAT_STRIDE;
  /* D is uniform across the wavefront: keep it in scalar registers */
  TC Dr[n1 * n1];
  {
    EP Dg = (EP)dr_tab<TC>(sd);
#pragma unroll
    for (int t = 0; t < n1 * n1; ++t)
      Dr[t] = uniform(Dg[t]);
  }
  const TC wi0 = (TC)sd.wi0;
  EP G = (EP)geo_tab<TC>(sd) + (size_t)e * GEO_STRIDE;
  cptr<int32_t> nbe = cglob(sd.nbr_elem) + (size_t)e * 6;
  /* face codes as vector byte loads: reading them as scalar words measured
   * slower (2.21 against 2.14 ms per 48^3 N=7 FP32 stage) */
  gptr<const uint8_t> nbc = glob(sd.nbr_code) + (size_t)e * 6;
  auto ncode = [&](int f) -> int { return (int)nbc[f]; };

  /* LDS word of node (i, j, k) */
  auto at3 = [&](int i, int j, int k) -> int {
    if constexpr (CF::swz)
      return S * k + n1 * (j ^ k) + (i ^ ((k & 1) << 2));
    else
      return i + n1 * j + S * k;
  };
  /* x-line (j, k) of an LDS field as whole vectors: with the swizzle its
   * halves i = 0..3 and 4..7 are the aligned 4-word groups at at3(0, j, k)
   * and at3(0, j, k) ^ 4 */
  using V4 = TL __attribute__((ext_vector_type(4)));
  auto xline_load = [&](const TL *base, int j, int k, TC *out) {
    if constexpr (CF::swz)
    {
      const int x = at3(0, j, k);
      const V4 lo = *reinterpret_cast<const V4 *>(base + x);
      const V4 hi = *reinterpret_cast<const V4 *>(base + (x ^ 4));
#pragma unroll
      for (int t = 0; t < 4; ++t)
      {
        out[t] = lo[t];
        out[4 + t] = hi[t];
      }
    }
    else
    {
#pragma unroll
      for (int i = 0; i < n1; ++i)
        out[i] = base[at3(i, j, k)];
    }
  };
  auto xline_store = [&](TL *base, int j, int k, const TC *in) {
    if constexpr (CF::swz)
    {
      const int x = at3(0, j, k);
      V4 lo, hi;
#pragma unroll
      for (int t = 0; t < 4; ++t)
      {
        lo[t] = in[t];
        hi[t] = in[4 + t];
      }
      *reinterpret_cast<V4 *>(base + x) = lo;
      *reinterpret_cast<V4 *>(base + (x ^ 4)) = hi;
    }
    else
    {
#pragma unroll
      for (int i = 0; i < n1; ++i)
        base[at3(i, j, k)] = in[i];
    }
  };

  /* ---- neighbour traces at face node (pa, pb) of both faces of axis A:
   * v1 v2 v3 and the stress column of A ---- */
  auto gather = [&](int A, TC (*nb)[6]) {
#pragma unroll
    for (int side = 0; side < 2; ++side)
    {
      const int f = 2 * A + side;
      const int32_t ep = nbe[f];
      const int code = ncode(f);
      if (code & NBR_GLUED)
      {
        /* flux fluctuation already on this face node (mortar kernel) */
        gptr<const TC> gb = glob(static_cast<const TC *>(sd.glue_buf)) + ep + pa + n1 * pb;
#pragma unroll
        for (int c = 0; c < 6; ++c)
          nb[side][c] = gb[c * n1 * n1];
        continue;
      }
      const int oi = perm_apply(nbr_perm(code), n1, pa, pb);
      const int np = face_node<3>(code & 7, n1, oi % n1, oi / n1);
      gptr<const TS> qp = qin + (size_t)ep * NF * Np + np;
#pragma unroll
      for (int c = 0; c < 3; ++c)
      {
        nb[side][c] = (TC)qp[c * Np];
        nb[side][3 + c] = (TC)qp[sfield(c, A) * Np];
      }
      if (code & NBR_BOUNDARY)
      {
        TC sv, ss;
        nbr_mirror<TC>(code, &sv, &ss);
#pragma unroll
        for (int c = 0; c < 3; ++c)
        {
          nb[side][c] *= sv;
          nb[side][3 + c] *= ss;
        }
      }
    }
  };

  /* ---- 1. stream the state into LDS with coalesced row reads: lane l
   * takes VW consecutive nodes at VW*(l + 64 r) of each field row of the
   * task's elements (flat over element and node) ---- */
  constexpr int VW = CoalVW<N, TS>::value;
  constexpr int ROWS = (CF::EPW * Np + 64 * VW - 1) / (64 * VW);
  const int nact = tk.ne * Np; /* valid (element, node) pairs of the task */
  auto coal_pos = [&](int r, int *lel, int *n) -> bool {
    const int q = VW * (lane + 64 * r);
    *lel = q / Np;
    *n = q - (q / Np) * Np;
    return q < nact;
  };
  auto lds_node = [&](int n) -> int {
    const int i = n % n1, j = (n / n1) % n1, k = n / (n1 * n1);
    return at3(i, j, k);
  };
  /* VW consecutive nodes n..n+VW-1 (n % VW == 0) stay on one x-line when
   * VW divides n1; they are then contiguous in LDS and move as one vector */
  constexpr bool vec_lds = (n1 % VW) == 0;
  using VL = TL __attribute__((ext_vector_type(VW)));
  using VC = TC __attribute__((ext_vector_type(VW)));
  if constexpr (dma)
  {
    if (!rows_ready)
    {
#pragma unroll
      for (int f = 0; f < NF; ++f)
        rows_to_lds<N, TS>(qin, tk.e0, tk.ne, f, reinterpret_cast<TS *>(lds) + f * FST);
    }
  }
  else
  {
    TS v[NF][ROWS][VW];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int r = 0; r < ROWS; ++r)
      {
        int lel, n;
        if (coal_pos(r, &lel, &n))
        {
          gptr<const TS> src = qin + ((size_t)(tk.e0 + lel) * NF + f) * Np + n;
#pragma unroll
          for (int t = 0; t < VW; ++t)
            v[f][r][t] = src[t];
        }
      }
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int r = 0; r < ROWS; ++r)
      {
        int lel, n;
        if (coal_pos(r, &lel, &n))
        {
          TL *dst = reinterpret_cast<TL *>(lds) + f * FST + lel * FS;
          if constexpr (vec_lds)
          {
            VL x;
#pragma unroll
            for (int t = 0; t < VW; ++t)
              x[t] = (TL)v[f][r][t];
            *reinterpret_cast<VL *>(dst + lds_node(n)) = x;
          }
          else
          {
#pragma unroll
            for (int t = 0; t < VW; ++t)
              dst[lds_node(n + t)] = (TL)v[f][r][t];
          }
        }
      }
  }
  const bool rhs_out = (mode == MODE_RHS || PART != PART_FUSED);
  const TR ca = (TR)cx.a, cdt = (TR)cx.dt;
  const TS cb = (TS)cx.b;
  constexpr int PD = sizeof(TS) == 4 ? NF : 3;
  /* stages with A = 0 (the first of each step) read no dq.  FP32 and FP64
   * store none: the old dq is rebuilt from the previous stage's state,
   * still in the other ping-pong buffer, B_{s-1} dq_{s-1} = q_s - q_{s-1}
   * (a holds A_s / B_{s-1}); these nodes are overwritten only below, by
   * this lane, after the read */
  const bool rd_dq = cx.a != 0.0;
  constexpr bool REC = sizeof(TS) == sizeof(TR);
  const bool rec = REC && cx.rec;
  const void *qprev = sd.q[cx.parity ^ 1];
  const void *dsrc = rec ? qprev : sd.dq;
  TR dqv[PD][ROWS][VW];
  TS qv[q_from_lds ? 1 : PD][ROWS][VW];
  auto fetch = [&](int f) {
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
    {
      int lel, n;
      if (!coal_pos(r, &lel, &n))
        continue;
      const size_t off = ((size_t)(tk.e0 + lel) * NF + f) * Np + n;
      gptr<const TR> dqe = glob(static_cast<const TR *>(dsrc)) + off;
#pragma unroll
      for (int t = 0; t < VW; ++t)
        dqv[f % PD][r][t] = rd_dq ? dqe[t] : TR(0);
      if constexpr (!q_from_lds)
      {
#pragma unroll
        for (int t = 0; t < VW; ++t)
          qv[f % PD][r][t] = qin[off + t];
      }
    }
  };
  /* the x- and y-face traces are requested with the state; each face
   * code's byte load makes them wait for the state rows to land, which
   * measured faster than issuing them alongside */
  TC nbx[2][6], nby[2][6], nbz[2][6];
  if (act && PART != PART_VOLUME)
  {
    gather(0, nbx);
    gather(1, nby);
  }

  TC rhoi = 0, lam = 0, mu = 0, Zpm = 0, Zsm = 0;
  if (act)
  {
    EP M = mtab + (size_t)e * MAT_STRIDE;
    rhoi = M[MAT_RHOI];
    lam = M[MAT_LAM];
    mu = M[MAT_MU];
    Zpm = M[MAT_ZP];
    Zsm = M[MAT_ZS];
  }
  const TC l2m = lam + TC(2) * mu;

  TC rate[NF][n1];
  if (PART == PART_SURFACE)
  {
    if (act)
    {
      gptr<const double> rh = glob((const double *)sd.rhs) + (size_t)e * NF * Np + n1 * L;
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int i = 0; i < n1; ++i)
          rate[f][i] = (TC)rh[f * Np + i];
    }
  }
  else
  {
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int i = 0; i < n1; ++i)
        rate[f][i] = TC(0);
  }

  /* LDS-DMA rows of this task have landed once vmcnt drains */
  if constexpr (dma)
    __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  /*
   * One axis pass.  The lane's line runs along axis A through (pa, pb); its
   * two ends are face node (pa, pb) of faces 2A and 2A+1.  For each of the
   * six raw quantities the line is contracted with D, the face lifts are
   * added at the ends, and the result is folded into the rates of the x-line
   * owners (directly for A = 0, through the transpose buffer otherwise).
   */
  auto pass = [&](auto AX, const TC (*nb)[6]) {
    constexpr int A = decltype(AX)::value;
    auto at = [&](int m) -> int {
      return A == 0 ? at3(m, pa, pb) : (A == 1 ? at3(pa, m, pb) : at3(pa, pb, m));
    };
    TC fl[2][6];
    if (act && PART != PART_VOLUME)
    {
#pragma unroll
      for (int side = 0; side < 2; ++side)
      {
        const int m = side ? N : 0;
        EP F = fmt + (2 * A + side) * FMAT_STRIDE;
        if (ncode(2 * A + side) & NBR_GLUED)
        {
#pragma unroll
          for (int c = 0; c < 6; ++c)
            fl[side][c] = nb[side][c];
          continue;
        }
        TC vm[3], Sm[3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
        {
          vm[c] = (TC)sq[c * FST + at(m)];
          Sm[c] = (TC)sq[sfield(c, A) * FST + at(m)];
        }
        riemann_axis<A, TC>(side ? TC(1) : TC(-1), vm, Sm, &nb[side][0],
                            &nb[side][3], Zpm, Zsm, F[FMAT_ZP], F[FMAT_ZS],
                            F[FMAT_IZP], F[FMAT_IZS], &fl[side][0], &fl[side][3]);
      }
    }
    const TC g = G[GEO_GX + A];
    const TC glam = g * lam, gl2m = g * l2m, gmu = g * mu;
#pragma unroll
    for (int c = 0; c < 6; ++c)
    {
      const int f = qfield<A>(c);
      TC q[n1];
      if (act)
      {
        TC w[n1];
        if constexpr (A == 0)
          xline_load(sq + f * FST, pa, pb, w);
        else
#pragma unroll
          for (int m = 0; m < n1; ++m)
            w[m] = (TC)sq[f * FST + at(m)];
#pragma unroll
        for (int j = 0; j < n1; ++j)
        {
          TC acc = 0;
          if (PART != PART_SURFACE)
#pragma unroll
            for (int m = 0; m < n1; ++m)
              acc += Dr[j + m * n1] * w[m];
          q[j] = acc;
        }
        /* lift (sJ/J)/w_0 = (2/h)/w_0: relative to the 2/h of the
         * derivative it is 1/w_0 */
        if (PART != PART_VOLUME)
        {
          const TC t0 = c < 3 ? fl[0][3 + c] : -fl[0][c == 3 ? A : tang(A, c - 4)];
          const TC t1 = c < 3 ? fl[1][3 + c] : fl[1][c == 3 ? A : tang(A, c - 4)];
          q[0] += wi0 * t0;
          q[N] += wi0 * t1;
        }
      }
      if constexpr (A == 0)
      {
        if (act)
#pragma unroll
          for (int i = 0; i < n1; ++i)
          {
            TC r9[NF];
#pragma unroll
            for (int ff = 0; ff < NF; ++ff)
              r9[ff] = rate[ff][i];
            fold<0, TC>(c, q[i], g, glam, gl2m, gmu, r9);
#pragma unroll
            for (int ff = 0; ff < NF; ++ff)
              rate[ff][i] = r9[ff];
          }
      }
      else
      {
        if (act)
#pragma unroll
          for (int j = 0; j < n1; ++j)
            tb[at(j)] = q[j];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        TC qx[n1];
        if (act)
          xline_load(tb, pa, pb, qx);
        if (act)
#pragma unroll
          for (int i = 0; i < n1; ++i)
          {
            const TC qv = qx[i];
            TC r9[NF];
#pragma unroll
            for (int ff = 0; ff < NF; ++ff)
              r9[ff] = rate[ff][i];
            fold<A, TC>(c, qv, g, glam, gl2m, gmu, r9);
#pragma unroll
            for (int ff = 0; ff < NF; ++ff)
              rate[ff][i] = r9[ff];
          }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
  };

  pass(std::integral_constant<int, 0>{}, nbx);
  /* sigma_11 is read by nothing after the x pass but the update: keep this
   * lane's row values of it and free its region for the transposes */
  const bool keep_s11 = q_from_lds && !(mode == MODE_RHS || PART != PART_FUSED);
  TL s11[ROWS][VW];
  if (keep_s11)
  {
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
    {
      int lel, n;
      if (!coal_pos(r, &lel, &n))
        continue;
      const TL *src = reinterpret_cast<TL *>(lds) + TBF * FST + lel * FS;
#pragma unroll
      for (int t = 0; t < VW; ++t)
        s11[r][t] = src[lds_node(n + t)];
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  /* the z-face traces load under the y pass */
  if (act && PART != PART_VOLUME)
    gather(2, nbz);
  pass(std::integral_constant<int, 1>{}, nby);
  pass(std::integral_constant<int, 2>{}, nbz);

#pragma unroll
  for (int i = 0; i < n1; ++i)
  {
    rate[0][i] *= rhoi;
    rate[1][i] *= rhoi;
    rate[2][i] *= rhoi;
  }

  /* ---- point source and manufactured forcing ---- */
  if (act && PART != PART_SURFACE && e == sd.src_elem)
  {
    gptr<const TC> w = src_tab<TC>(sd) + n1 * L;
    const TC g = (TC)cx.g;
#pragma unroll
    for (int f = 0; f < NF; ++f)
    {
      const TC ag = (TC)sd.src_amp[f] * g;
#pragma unroll
      for (int i = 0; i < n1; ++i)
        rate[f][i] += w[i] * ag;
    }
  }
  if (act && PART != PART_SURFACE && sd.mms)
  {
    gptr<const TC> T = mms_tab<TC>(sd) + (size_t)e * 6 * n1;
    const TC sw = (TC)cx.sw, cw = (TC)cx.cw;
    const TC Sy = T[2 * n1 + pa], Cy = T[3 * n1 + pa];
    const TC Sz = T[4 * n1 + pb], Cz = T[5 * n1 + pb];
#pragma unroll
    for (int i = 0; i < n1; ++i)
    {
      const TC Sv[3] = {T[i], Sy, Sz};
      const TC Cv[3] = {T[n1 + i], Cy, Cz};
#pragma unroll
      for (int f = 0; f < NF; ++f)
        rate[f][i] += (TC)sd.mms_F[f] * mms_pattern<TC>(f, Sv, Cv) * (f < 3 ? sw : cw);
    }
  }

  /* ---- 5. output, one field at a time: the line's rates go through the
   * transpose buffer so that dq and q are read and written as coalesced
   * rows (same lane -> node assignment as the load).  The dq (and, when the
   * LDS copy is rounded, q_old) rows are fetched PD fields ahead so that the
   * per-field LDS round trips do not wait on global memory ---- */
  /* fetching the dq rows before the passes instead (FP32: all nine fit in
   * registers) measured slower, 2.28 against 2.14 ms per 48^3 N=7 stage */
  if (!rhs_out)
  {
#pragma unroll
    for (int f = 0; f < PD; ++f)
      fetch(f);
  }
  const TC *tb0 = reinterpret_cast<TC *>(reinterpret_cast<TL *>(lds) + TBF * FST);
#pragma unroll
  for (int f = 0; f < NF; ++f)
  {
    if (act)
      xline_store(tb, pa, pb, rate[f]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    TC rv[ROWS][VW];
    TL qo_l[ROWS][VW];
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
    {
      int lel, n;
      if (!coal_pos(r, &lel, &n))
        continue;
      const TC *tbe = tb0 + lel * FS;
      const TL *sqe = reinterpret_cast<TL *>(lds) + f * FST + lel * FS;
      if constexpr (vec_lds)
      {
        const VC x = *reinterpret_cast<const VC *>(tbe + lds_node(n));
#pragma unroll
        for (int t = 0; t < VW; ++t)
          rv[r][t] = x[t];
        if constexpr (q_from_lds)
        {
          if (f == TBF)
          {
#pragma unroll
            for (int t = 0; t < VW; ++t)
              qo_l[r][t] = s11[r][t];
          }
          else
          {
            const VL y = *reinterpret_cast<const VL *>(sqe + lds_node(n));
#pragma unroll
            for (int t = 0; t < VW; ++t)
              qo_l[r][t] = y[t];
          }
        }
      }
      else
      {
#pragma unroll
        for (int t = 0; t < VW; ++t)
        {
          rv[r][t] = tbe[lds_node(n + t)];
          if constexpr (q_from_lds)
            qo_l[r][t] = f == TBF ? s11[r][t] : sqe[lds_node(n + t)];
        }
      }
    }
    /* every lane has its rows of this field: the buffer may be refilled */
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
    {
      int lel, n;
      if (!coal_pos(r, &lel, &n))
        continue;
      const size_t off = ((size_t)(tk.e0 + lel) * NF + f) * Np + n;
      if (rhs_out)
      {
        gptr<double> rh = glob(sd.rhs) + off;
#pragma unroll
        for (int t = 0; t < VW; ++t)
          rh[t] = (double)rv[r][t];
        continue;
      }
      gptr<TR> dqe = glob(static_cast<TR *>(sd.dq)) + off;
      gptr<TS> qo = glob(static_cast<TS *>(sd.q[cx.parity ^ 1])) + off;
#pragma unroll
      for (int t = 0; t < VW; ++t)
      {
        const TS qold = q_from_lds ? (TS)qo_l[r][t] : qv[q_from_lds ? 0 : f % PD][r][t];
        TR dqo = dqv[f % PD][r][t];
        if constexpr (REC)
          if (rec && rd_dq)
            dqo = (TR)qold - dqo;
        const TR d = ca * dqo + cdt * (TR)rv[r][t];
        if (!cx.dq_dead && !rec)
          dqe[t] = d;
        qo[t] = qold + cb * (TS)d;
      }
    }
    if (!rhs_out && f + PD < NF)
      fetch(f + PD);
    /* field f of this task is written: its LDS region takes the next
     * task's field f (its reads above have completed, their results being
     * stored) */
    if constexpr (dma)
    {
      if (nx.valid)
        rows_to_lds<N, TS>(nx.q, nx.e0, nx.ne, f, reinterpret_cast<TS *>(lds) + f * FST);
    }
  }
}

/* one task (EPW elements) per wavefront; one-wave workgroups so that LDS,
 * not the workgroup shape, bounds the wavefronts resident per CU */
constexpr int LINE_WPB = 1;

__device__ __forceinline__ int first_wave_task(const LaunchArgs &a, int *stride,
                                               int *end)
{
  /* the wave index is uniform in the wave; saying so keeps the task walk,
   * and every table read that follows from it, in scalar registers */
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int b = blockIdx.x, G = gridDim.x;
  if (a.xcd_remap != 1 || G % a.nxcd)
  {
    *stride = G * LINE_WPB;
    *end = a.ntasks;
    return b * LINE_WPB + w;
  }
  /* XCD x owns the contiguous task slice [x*C, (x+1)*C) */
  const int x = b % a.nxcd, s = b / a.nxcd, per = G / a.nxcd;
  const int C = (a.ntasks + a.nxcd - 1) / a.nxcd;
  const int lo = x * C;
  int hi = lo + C;
  if (hi > a.ntasks)
    hi = a.ntasks;
  *stride = per * LINE_WPB;
  *end = hi;
  return lo + s * LINE_WPB + w;
}

/* wavefronts per SIMD the register allocation of the line kernel aims at
 * (amdgpu_waves_per_eu; 1: no limit).  FP32 at N = 4 and 6 and FP32 rates
 * at N = 4 need 177-264 registers unconstrained, one or two wavefronts
 * per SIMD; capped for one more wavefront they spill 28-60 bytes per lane
 * and run faster (124 M DOF bricks, ms per stage: FP32 N=4 0.658 -> 0.573,
 * N=6 0.950 -> 0.618; FP32 rates N=4 2.03 -> 1.69).  Elsewhere the cap
 * measured slower (FP32 N=3 0.454 -> 0.604, N=5 0.566 -> 0.580, FP64 N=4
 * 1.18 -> 1.93) or is moot (LDS already limits N=7 and FP64 N=3 to two). */
template <int N, class P> struct LineWaves
{
  static constexpr int value = __is_same(P, P32) ? (N == 4 ? 3 : (N == 6 ? 2 : 1))
                               : __is_same(P, P32R) ? (N == 4 ? 3 : 1)
                                                    : 1;
};

template <int N, class P, int PART>
__global__ __launch_bounds__(64 * LINE_WPB)
__attribute__((amdgpu_waves_per_eu(LineWaves<N, P>::value))) void elastic_line_kernel(LaunchArgs a)
{
  using CF = LineCfg<N, typename P::TC, typename P::TC>;
  __shared__ __align__(16) char lds[LINE_WPB * CF::bytes];
  char *mine = lds + (threadIdx.x >> 6) * CF::bytes;
  using TS = typename P::TS;
  int stride, end;
  bool ready = false;
  for (int ti = first_wave_task(a, &stride, &end); ti < end; ti += stride)
  {
    const Task tk = load_task(a.tasks, ti);
    const auto &sd = *(cglob(a.subs) + tk.sub);
    const StageCtx cx = stage_ctx(sd, a);
    NextRows<TS> nx;
    nx.valid = ti + stride < end;
    nx.q = glob(static_cast<const TS *>(sd.q[cx.parity]));
    nx.e0 = nx.ne = 0;
    if (nx.valid)
    {
      const Task nt = load_task(a.tasks, ti + stride);
      const auto &ns = *(cglob(a.subs) + nt.sub);
      nx.q = glob(static_cast<const TS *>(ns.q[stage_ctx(ns, a).parity]));
      nx.e0 = nt.e0;
      nx.ne = nt.ne;
    }
    elastic_line_task<N, P, PART>(sd, tk, cx, a.mode, mine, ready, nx);
    ready = nx.valid;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

constexpr int lmax(int a, int b) { return a > b ? a : b; }

/* every order 3..7 in one launch: each wavefront switches on its task's N */
template <class P>
__global__ __launch_bounds__(64 * LINE_WPB) void elastic_line_mixed(LaunchArgs a)
{
  using TS = typename P::TS;
  using TC = typename P::TC;
  constexpr int W = lmax(
      lmax(lmax((int)LineCfg<3, TC, TC>::bytes, (int)LineCfg<4, TC, TC>::bytes),
           lmax((int)LineCfg<5, TC, TC>::bytes, (int)LineCfg<6, TC, TC>::bytes)),
      (int)LineCfg<7, TC, TC>::bytes);
  __shared__ __align__(16) char lds[LINE_WPB * W];
  char *mine = lds + (threadIdx.x >> 6) * W;
  int stride, end;
  for (int ti = first_wave_task(a, &stride, &end); ti < end; ti += stride)
  {
    const Task tk = load_task(a.tasks, ti);
    const auto &sd = *(cglob(a.subs) + tk.sub);
    const StageCtx cx = stage_ctx(sd, a);
    NextRows<TS> nx; /* no cross-task prefetch: the next task's order differs */
    nx.valid = false;
    nx.q = glob(static_cast<const TS *>(sd.q[cx.parity]));
    nx.e0 = nx.ne = 0;
    switch (tk.N)
    {
    case 3: elastic_line_task<3, P>(sd, tk, cx, a.mode, mine, false, nx); break;
    case 4: elastic_line_task<4, P>(sd, tk, cx, a.mode, mine, false, nx); break;
    case 5: elastic_line_task<5, P>(sd, tk, cx, a.mode, mine, false, nx); break;
    case 6: elastic_line_task<6, P>(sd, tk, cx, a.mode, mine, false, nx); break;
    default: elastic_line_task<7, P>(sd, tk, cx, a.mode, mine, false, nx); break;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

template <class P>
static hipError_t launch_line_n(int N, const LaunchArgs &a, int grid,
                                hipStream_t s, unsigned dyn)
{
  switch (N)
  {
#define BFAMX_CASE(NN)                                                         \
  case NN:                                                                     \
    hipLaunchKernelGGL((elastic_line_kernel<NN, P, PART_FUSED>), dim3(grid),   \
                       dim3(64 * LINE_WPB), dyn, s, a);                        \
    break;
    BFAMX_CASE(3)
    BFAMX_CASE(4)
    BFAMX_CASE(5)
    BFAMX_CASE(6)
    BFAMX_CASE(7)
#undef BFAMX_CASE
  default:
    hipLaunchKernelGGL((elastic_line_mixed<P>), dim3(grid), dim3(64 * LINE_WPB),
                       0, s, a);
    break;
  }
  return hipGetLastError();
}

template <class P>
static hipError_t occupancy_n(int N, int *blocks, unsigned dyn)
{
  switch (N)
  {
#define BFAMX_CASE(NN)                                                         \
  case NN:                                                                     \
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(                       \
        blocks, elastic_line_kernel<NN, P, PART_FUSED>, 64 * LINE_WPB, dyn);
    BFAMX_CASE(3)
    BFAMX_CASE(4)
    BFAMX_CASE(5)
    BFAMX_CASE(6)
    BFAMX_CASE(7)
#undef BFAMX_CASE
  default:
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, elastic_line_mixed<P>,
                                                        64 * LINE_WPB, dyn);
  }
}

hipError_t line_blocks_per_cu(int prec, int N, int *blocks, unsigned dyn)
{
  switch (prec)
  {
  case PREC_FP32_BF16X3: return hipErrorInvalidValue;
  case PREC_FP32: return occupancy_n<P32>(N, blocks, dyn);
  case PREC_FP32_RHS: return occupancy_n<P32R>(N, blocks, dyn);
  default: return occupancy_n<P64>(N, blocks, dyn);
  }
}

hipError_t launch_elastic_line(int prec, int N, const LaunchArgs &a, int grid,
                               hipStream_t s, unsigned dyn)
{
  switch (prec)
  {
  case PREC_FP32_BF16X3: return hipErrorInvalidValue; /* node-parallel kernel only */
  case PREC_FP32: return launch_line_n<P32>(N, a, grid, s, dyn);
  case PREC_FP32_RHS: return launch_line_n<P32R>(N, a, grid, s, dyn);
  default: return launch_line_n<P64>(N, a, grid, s, dyn);
  }
}

} // namespace bfamx
