/*
 * LSRK coefficient tables and the PCG32 generator.
 *
 * The coefficients are Carpenter & Kennedy's published rationals (NASA
 * TM-109112, 1994; KC54 is their (5,4) 2N-storage scheme), Williamson's
 * (J. Comput. Phys. 35, 1980) W33, Heun and forward Euler, evaluated in long
 * double and rounded to double exactly as bfamo_lsrk_init does
 * (bfam.c:8996-9090) so the FP64 tables agree bit for bit.  PCG32 is
 * O'Neill's XSH-RR 64/32 generator (pcg-random.org, 2014), the generator
 * behind bfam_pcg32_* (bfam.c:8328-8387).
 */
#include "../bfamx_internal.h"

int bfamx_lsrk_init(bfamx_lsrk_t *ts, bfamx_lsrk_method_t method)
{
  if (!ts)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_lsrk_init: NULL table");
  for (int i = 0; i < 8; ++i)
    ts->A[i] = ts->B[i] = ts->C[i] = 0;
  ts->C[8] = 0;
  switch (method)
  {
  case BFAMX_LSRK_FE:
    ts->n_stages = 1;
    ts->B[0] = 1.0;
    ts->C[1] = 1.0;
    break;
  case BFAMX_LSRK_HEUN:
    ts->n_stages = 2;
    ts->A[1] = -1.0;
    ts->B[0] = 1.0;
    ts->B[1] = 0.5;
    ts->C[1] = 1.0;
    ts->C[2] = 1.0;
    break;
  case BFAMX_LSRK_W33:
    ts->n_stages = 3;
    ts->A[1] = (double)(-5.0L / 9.0L);
    ts->A[2] = (double)(-153.0L / 128.0L);
    ts->B[0] = (double)(1.0L / 3.0L);
    ts->B[1] = (double)(15.0L / 16.0L);
    ts->B[2] = (double)(8.0L / 15.0L);
    ts->C[1] = (double)(1.0L / 3.0L);
    ts->C[2] = (double)(3.0L / 4.0L);
    ts->C[3] = 1.0;
    break;
  case BFAMX_LSRK_KC54:
  default:
    ts->n_stages = 5;
    ts->A[1] = (double)(-567301805773.0L / 1357537059087.0L);
    ts->A[2] = (double)(-2404267990393.0L / 2016746695238.0L);
    ts->A[3] = (double)(-3550918686646.0L / 2091501179385.0L);
    ts->A[4] = (double)(-1275806237668.0L / 842570457699.0L);
    ts->B[0] = (double)(1432997174477.0L / 9575080441755.0L);
    ts->B[1] = (double)(5161836677717.0L / 13612068292357.0L);
    ts->B[2] = (double)(1720146321549.0L / 2090206949498.0L);
    ts->B[3] = (double)(3134564353537.0L / 4481467310338.0L);
    ts->B[4] = (double)(2277821191437.0L / 14882151754819.0L);
    ts->C[1] = ts->B[0];
    ts->C[2] = (double)(2526269341429.0L / 6820363962896.0L);
    ts->C[3] = (double)(2006345519317.0L / 3224310063776.0L);
    ts->C[4] = (double)(2802321613138.0L / 2924317926251.0L);
    ts->C[5] = 1.0;
    break;
  }
  return BFAMX_OK;
}

uint32_t bfamx_pcg32_next(bfamx_pcg32_t *rng)
{
  const uint64_t s = rng->state;
  rng->state = s * 6364136223846793005ULL + (rng->inc | 1u);
  const uint32_t x = (uint32_t)(((s >> 18) ^ s) >> 27);
  const uint32_t r = (uint32_t)(s >> 59);
  return (x >> r) | (x << ((32u - r) & 31u));
}

void bfamx_pcg32_seed(bfamx_pcg32_t *rng, uint64_t initstate, uint64_t initseq)
{
  rng->state = 0;
  rng->inc = (initseq << 1) | 1u;
  (void)bfamx_pcg32_next(rng);
  rng->state += initstate;
  (void)bfamx_pcg32_next(rng);
}

uint32_t bfamx_pcg32_bounded(bfamx_pcg32_t *rng, uint32_t bound)
{
  if (bound == 0)
    return 0;
  const uint32_t floor_ = (uint32_t)(0x100000000ULL % bound);
  for (;;)
  {
    const uint32_t r = bfamx_pcg32_next(rng);
    if (r >= floor_)
      return r % bound;
  }
}

double bfamx_pcg32_uniform(bfamx_pcg32_t *rng)
{
  return bfamx_pcg32_next(rng) * (1.0 / 4294967296.0);
}
