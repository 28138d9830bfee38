/*
 * Derivative contractions of the node-parallel stage on the matrix cores,
 * in split-BF16x3: every FP32 operand x is split into x_hi = bf16(x) and
 * x_lo = bf16(x - x_hi) (round to nearest even), and a product is taken as
 * a_hi b_hi + a_hi b_lo + a_lo b_hi with FP32 accumulation (the a_lo b_lo
 * term, about 2^-16 of the product, is dropped).  Each operand then carries
 * about 16 significant bits instead of FP32's 24.
 *
 * One axis of one workgroup's elements at a time: the lines along the axis
 * (n1^2 per element and field, for the six fields the axis contributes to)
 * are the 16 rows of an MFMA tile, the line's N+1 nodes are K (padded with
 * zeros to 16), and the output node index i is the tile's column (padded
 * to 16):  C[line][i] = sum_a u[line][a] D[i][a], three
 * v_mfma_f32_16x16x16_bf16 per tile.  Lane layout of that instruction:
 * A[m][k] at lane m + 16 (k / 4), item k % 4; B[k][n] at lane n + 16 (k / 4),
 * item k % 4; C[m][n] at lane n + 16 (m / 4), item m % 4.  Only full
 * wavefronts take tiles, so every lane of an MFMA holds defined operands.
 */
#pragma once

#include "common.hpp"

namespace bfamx
{

typedef short mf_bf16x4 __attribute__((ext_vector_type(4)));
typedef float mf_f32x4 __attribute__((ext_vector_type(4)));

/* the bf16 nearest to x (ties to even), as its 16 bits */
__device__ __forceinline__ short mf_bf16(float x)
{
  unsigned u = __builtin_bit_cast(unsigned, x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float mf_float(short h)
{
  return __builtin_bit_cast(float, (unsigned)(unsigned short)h << 16);
}
/* x = hi + lo to about 2^-17 |x| */
__device__ __forceinline__ void mf_split(float x, short &hi, short &lo)
{
  hi = mf_bf16(x);
  lo = mf_bf16(x - mf_float(hi));
}

/* volume node of point a on line (p, q) along axis AX */
template <int AX, int n1>
__device__ __forceinline__ int mf_node(int p, int q, int a)
{
  return AX == 0 ? a + n1 * (p + n1 * q) : (AX == 1 ? p + n1 * (a + n1 * q) : p + n1 * (q + n1 * a));
}

/*
 * out[s * EPB * Np + le * Np + node] = the derivative along axis AX (not yet
 * scaled by 2/h) of field fld[s] of element le, for the EPB elements staged
 * in sq ([EPB][9][Np] FP32) and D in sD (column-major, n1 x n1).  Called by
 * every thread of the block; the caller synchronises before reading out.
 */
template <int N, int EPB, int AX>
__device__ __forceinline__ void mf_axis(const float *sq, const float *sD, float *out,
                                        const int (&fld)[6])
{
  constexpr int n1 = N + 1, Np = n1 * n1 * n1, NL = n1 * n1;
  constexpr int L = EPB * 6 * NL;    /* lines */
  constexpr int NT = (L + 15) / 16;  /* tiles */
  constexpr int NFULL = (EPB * Np) / 64; /* full wavefronts of the block */
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (wave >= NFULL)
    return;
  const int lane = threadIdx.x & 63, col = lane & 15, kb = lane >> 4;
  mf_bf16x4 bh, bl;
#pragma unroll
  for (int t = 0; t < 4; ++t)
  {
    const int a = 4 * kb + t;
    const float d = (col < n1 && a < n1) ? sD[col + a * n1] : 0.f;
    short h, l;
    mf_split(d, h, l);
    bh[t] = h;
    bl[t] = l;
  }
  for (int tile = wave; tile < NT; tile += NFULL)
  {
    mf_bf16x4 ah, al;
    {
      const int line = tile * 16 + col;
      const int le = line / (6 * NL), rem = line - le * 6 * NL;
      const int s = rem / NL, pq = rem - s * NL;
      const float *u = sq + (le * 9 + fld[s < 6 ? s : 0]) * Np;
#pragma unroll
      for (int t = 0; t < 4; ++t)
      {
        const int a = 4 * kb + t;
        const float x =
            (line < L && a < n1) ? u[mf_node<AX, n1>(pq % n1, pq / n1, a)] : 0.f;
        short h, l;
        mf_split(x, h, l);
        ah[t] = h;
        al[t] = l;
      }
    }
    mf_f32x4 c = {0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, c, 0, 0, 0);
    if (col < n1)
    {
#pragma unroll
      for (int t = 0; t < 4; ++t)
      {
        const int line = tile * 16 + 4 * kb + t;
        if (line < L)
        {
          const int le = line / (6 * NL), rem = line - le * 6 * NL;
          const int s = rem / NL, pq = rem - s * NL;
          out[(s * EPB + le) * Np + mf_node<AX, n1>(pq % n1, pq / n1, col)] = c[t];
        }
      }
    }
  }
}

} // namespace bfamx
