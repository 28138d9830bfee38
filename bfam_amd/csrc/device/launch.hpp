/* Host-side entry points of the kernel translation units (C++ linkage, used
 * only by runtime.hip). */
#pragma once

#include "common.hpp"

namespace bfamx
{

/* elements per workgroup for order N (threads = epb * Np, at most 512) */
inline int elastic_epb(int N)
{
  const int np = (N + 1) * (N + 1) * (N + 1);
  const int e = 512 / np;
  return e < 1 ? 1 : e;
}
/* line-formulation elastic stage: elements per task (one wavefront) */
inline int elastic_line_epw(int N)
{
  const int nl = (N + 1) * (N + 1);
  return 64 / nl;
}
/* workgroups of the line stage for ntasks tasks (one task per wave) */
constexpr int LINE_TASKS_PER_BLOCK = 1;
/* resident one-wave blocks per CU of the elastic stage kernel, with dyn
 * bytes of dynamic LDS per block */
hipError_t line_blocks_per_cu(int prec, int N, int *blocks, unsigned dyn = 0);
/* dyn: dynamic LDS bytes per block that the kernel does not use, requested
 * only to cap how many blocks share a CU */
hipError_t launch_elastic_line(int prec, int N, const LaunchArgs &a, int grid,
                               hipStream_t s, unsigned dyn = 0);

inline int acoustic_epb(int N)
{
  const int np = (N + 1) * (N + 1);
  const int e = 256 / np;
  return e < 1 ? 1 : e;
}

/* precision codes as in bfamx_precision_t */
enum
{
  PREC_FP64 = 0,
  PREC_FP64_PLAIN = 1,
  PREC_FP32_RHS = 2,
  PREC_FP32 = 3,
  PREC_FP32_BF16X3 = 4
};
constexpr int NPREC = 5;

/* fused stage or RHS-only launch over a task table.  N < 0: mixed orders
 * (one kernel handles every N in 3..7).  grid: number of workgroups. */
hipError_t launch_elastic_stage(int prec, int N, const LaunchArgs &a, int grid,
                                hipStream_t s, unsigned dyn = 0, int alt = 0);
/* resident blocks per CU of the node-parallel stage kernel (N = 3..7);
 * alt: with its other block size (kernel kinds 3, 4) */
hipError_t node_blocks_per_cu(int prec, int N, int *blocks, unsigned dyn = 0, int alt = 0);
hipError_t launch_acoustic_stage(int prec, int N, const LaunchArgs &a,
                                 int grid, hipStream_t s);

/* unfused FP64 path: volume kernel (rhs = volume + sources), surface kernel
 * (rhs += lifted fluctuations), then the element-wise LSRK update */
hipError_t launch_elastic_plain(int N, const LaunchArgs &a, int grid,
                                hipStream_t s);
hipError_t launch_acoustic_plain(int N, const LaunchArgs &a, int grid,
                                 hipStream_t s);
hipError_t launch_plain_update(const SubDev *sd_dev, size_t ndof,
                               hipStream_t s, int nblocks);
/* advance the clocks of nsubs device descriptors by one stage */
hipError_t launch_clock_advance(const SubDev *subs, int nsubs, hipStream_t s);
hipError_t launch_mortar(int prec, const GlueArgs &a, hipStream_t s);

/* field transfer between adapted meshes: one source or destination
 * subdomain (state of the current parity, order, elements) */
struct XferSide
{
  void *q;
  int N;
  int K;
};
/* one destination element and its 1 (same / refine) or 8 (coarsen)
 * sources; bits 4d..4d+3 of op[c]: which 1-D operator of the order pair
 * (0 Pr, 1 Pr_cfb, 2 Pr_cft, 3 Pr_r2b, 4 Pr_r2t) source c takes along axis
 * d.  Every member is a 32-bit word, so that the kernel reads the record
 * with aligned loads only (packed bytes were read two at a time at odd
 * addresses) */
struct XferTask
{
  int32_t dst_sub, dst_elem;
  int32_t nsrc;
  int32_t src_sub[8], src_elem[8];
  int32_t op[8];
};
/* ops: the FP64 table of glue_ops_index (kind 0) */
hipError_t launch_transfer(int src_fp32, int dst_fp32, const XferSide *srcs,
                           const XferSide *dsts, const XferTask *tasks, int ntasks,
                           const double *ops, hipStream_t s);

/* reductions: per-element partial sums (FP64) of the energy density (K), or
 * interleaved pairs (2K) of |qa - qb|^2, |qb|^2 and of the error against the
 * manufactured solution with time factors cos/sin(omega t) */
hipError_t launch_cfl(const SubDev &sd, double *partial, hipStream_t s);
hipError_t launch_energy(const SubDev &sd, int state_is_fp32, int parity,
                         double *partial, hipStream_t s);
hipError_t launch_l2diff(const SubDev &a, int a_fp32, int pa, const SubDev &b,
                         int b_fp32, int pb, double *partial, hipStream_t s);
hipError_t launch_mms_error(const SubDev &sd, int state_is_fp32, int parity,
                            const double *A, double cwt, double swt,
                            double *partial, hipStream_t s);

} // namespace bfamx
