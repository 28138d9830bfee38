/*
 * Field transfer between two adapted meshes (h-refine, h-coarsen, order
 * change), run once between steps after the mesh changes.
 *
 * The reference supplies the transfer maps (bfam_domain_pxest_transfer_
 * maps_init, bfam.c:5205-5387) and uploads, per order pair, the 1-D
 * operators Pr, Pr_cfb, Pr_cft, Pr_r2b, Pr_r2t (bfam.c:8399-8404, from
 * create_interpolators, bfam.c:5606); the kernel that applies them was
 * application code.  Here one workgroup builds one destination element:
 * for each of its sources (one, or the eight children of a coarsened
 * element) and each field, the source element's values pass through LDS
 * three times, one 1-D operator along x, then y, then z (sum
 * factorisation: (N_d+1)(N_s+1)^3 + (N_d+1)^2 (N_s+1)^2 + (N_d+1)^3 (N_s+1)
 * multiply-adds instead of (N_d+1)^3 (N_s+1)^3), and the result adds into
 * the element's accumulator in LDS.  Arithmetic is FP64 whatever the state
 * types; each destination value is rounded once to its state type.  It is
 * HBM-bound and runs once per adaptation, so it is written for clarity:
 * 256 threads, node-strided loops.
 */
#include "launch.hpp"

namespace bfamx
{

constexpr int XFER_THREADS = 256;
constexpr int XFER_MAXNP = 512; /* (7+1)^3: elastic orders 3..7 */

template <class TS, class TD>
__global__ void __launch_bounds__(XFER_THREADS)
    transfer_kernel(const XferSide *__restrict__ srcs, const XferSide *__restrict__ dsts,
                    const XferTask *__restrict__ tasks, const double *__restrict__ ops)
{
  __shared__ double buf0[XFER_MAXNP], buf1[XFER_MAXNP];
  __shared__ double acc[9 * XFER_MAXNP];
  const XferTask &t = tasks[blockIdx.x];
  const XferSide &dd = dsts[t.dst_sub];
  const int nd = dd.N + 1, npd = nd * nd * nd;
  for (int i = threadIdx.x; i < 9 * npd; i += XFER_THREADS)
    acc[i] = 0.0;
  for (int c = 0; c < t.nsrc; ++c)
  {
    const XferSide &ss = srcs[t.src_sub[c]];
    const int ns = ss.N + 1, nps = ns * ns * ns;
    const int op = t.op[c];
    const double *Px = ops + glue_ops_index(0, ss.N, dd.N, op & 15);
    const double *Py = ops + glue_ops_index(0, ss.N, dd.N, (op >> 4) & 15);
    const double *Pz = ops + glue_ops_index(0, ss.N, dd.N, (op >> 8) & 15);
    const TS *q = static_cast<const TS *>(ss.q) + (size_t)t.src_elem[c] * 9 * nps;
    for (int f = 0; f < 9; ++f)
    {
      __syncthreads(); /* buf0 free (previous field's z pass done) */
      for (int n = threadIdx.x; n < nps; n += XFER_THREADS)
        buf0[n] = (double)q[f * nps + n];
      __syncthreads();
      /* x: buf1[i + nd (j + ns k)], i < nd, j, k < ns */
      for (int n = threadIdx.x; n < nd * ns * ns; n += XFER_THREADS)
      {
        const int i = n % nd, jk = n / nd;
        double s = 0.0;
        for (int a = 0; a < ns; ++a)
          s += Px[i + nd * a] * buf0[a + ns * jk];
        buf1[n] = s;
      }
      __syncthreads();
      /* y: buf0[i + nd (j + nd k)], i, j < nd, k < ns */
      for (int n = threadIdx.x; n < nd * nd * ns; n += XFER_THREADS)
      {
        const int i = n % nd, j = (n / nd) % nd, k = n / (nd * nd);
        double s = 0.0;
        for (int a = 0; a < ns; ++a)
          s += Py[j + nd * a] * buf1[i + nd * (a + ns * k)];
        buf0[n] = s;
      }
      __syncthreads();
      /* z: acc[f][i + nd (j + nd k)], i, j, k < nd */
      for (int n = threadIdx.x; n < npd; n += XFER_THREADS)
      {
        const int ij = n % (nd * nd), k = n / (nd * nd);
        double s = 0.0;
        for (int a = 0; a < ns; ++a)
          s += Pz[k + nd * a] * buf0[ij + nd * nd * a];
        acc[f * npd + n] += s;
      }
    }
  }
  __syncthreads();
  TD *out = static_cast<TD *>(dd.q) + (size_t)t.dst_elem * 9 * npd;
  for (int i = threadIdx.x; i < 9 * npd; i += XFER_THREADS)
    out[i] = (TD)acc[i];
}

hipError_t launch_transfer(int src_fp32, int dst_fp32, const XferSide *srcs,
                           const XferSide *dsts, const XferTask *tasks, int ntasks,
                           const double *ops, hipStream_t s)
{
  if (ntasks <= 0)
    return hipSuccess;
  const dim3 grid(ntasks), block(XFER_THREADS);
  if (src_fp32 && dst_fp32)
    transfer_kernel<float, float><<<grid, block, 0, s>>>(srcs, dsts, tasks, ops);
  else if (src_fp32)
    transfer_kernel<float, double><<<grid, block, 0, s>>>(srcs, dsts, tasks, ops);
  else if (dst_fp32)
    transfer_kernel<double, float><<<grid, block, 0, s>>>(srcs, dsts, tasks, ops);
  else
    transfer_kernel<double, double><<<grid, block, 0, s>>>(srcs, dsts, tasks, ops);
  return hipGetLastError();
}

} // namespace bfamx
