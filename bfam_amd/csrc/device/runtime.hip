/*
 * bfamx device runtime: the C ABI of include/bfamx.h on top of HIP.
 *
 * Replaces bfam's OCCA device layer (bfamo_*, bfam.c:8389-9092): device
 * creation with an HBM budget (bfamo_device_malloc's mem_fudge check,
 * bfam.c:8561-8589, returning BFAMX_EBUDGET instead of aborting), upload of
 * the per-N operators and the integer maps (bfamo_add_operator_*,
 * bfam.c:8712-8828), and the per-stage launches that the reference's
 * applications issued through BFAMO_KERNEL_RUN (bfam.h:1673-1719).  On one
 * GPU the glue trace exchange (bfamo_communicator_post_* and _wait,
 * bfam.c:8900-8929) is a gather inside the stage kernel.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../bfamx_internal.h"
#include "launch.hpp"

using namespace bfamx;

struct bfamx_device
{
  int hip_dev;
  hipStream_t stream;
  size_t budget;
  size_t used;
  int num_cu;
  int num_xcd;
  int xcd_remap;
  hipEvent_t ev0, ev1;
  /* mixed-order launch strategy chosen by bfamx_batch_autotune, per
   * precision and set of orders present (bit N): -1 unknown, 0 one mixed
   * launch, 1 one launch per order */
  signed char mixed_choice[NPREC][256];
  /* elastic stage launches with at most the resident wavefronts, each
   * looping over tasks and loading the next task's state during the
   * current task's output.  Off by default: on gfx950 the wait for the
   * prefetched rows also waits for the previous task's stores (one vmcnt for
   * loads and stores), and the FP32 48^3 N=7 stage measured 2.75 ms against
   * 2.34 ms with one wavefront per task */
  int persistent;
  int line_bpc[NPREC][9]; /* cached blocks per CU, 0 unknown */
  /* elastic stage kernel per precision and order, chosen by
   * bfamx_batch_autotune: -1 unknown (the default of the per-order
   * measurements), 0 line (one wavefront per element, lanes own lines), 1
   * node-parallel (threads own nodes), 2 node-parallel on a resident grid
   * (as many blocks as fit the CUs at once, each looping over tasks), 3 and
   * 4 as 1 and 2 with the kernel's other block size (N = 3..5) */
  signed char kchoice[NPREC][9];
  /* resident-block cap per precision, order and kernel slot (0 line, 1
   * node-parallel, 2 node-parallel with its other block size), chosen by
   * bfamx_batch_autotune or bfamx_set_occupancy:
   * 0 as many blocks as registers and LDS allow, 1 three quarters of that,
   * 2 half.  A cap is a request of dynamic LDS the kernel does not use
   * (occ_dyn, bytes per block, found once by the occupancy calculator;
   * ~0u unknown) */
  signed char occ[NPREC][9][3];
  unsigned occ_dyn[NPREC][9][3][3];
  /* the per-order launches of a grouped batch run on side streams forked
   * from and joined back to `stream` (parallel branches when captured) */
  hipStream_t side[8];
  hipEvent_t fork_ev, join_ev[8];
  /* device address of this GPU's copy of `scratch` (bfamx_rhs_noalloc) */
  double *scratch;
};

/* Staging buffer of bfamx_rhs_noalloc: one copy per GPU, held until the
 * process ends.  Host code uses only the per-device address that
 * hipGetSymbolAddress returns (bfamx_device::scratch), never this name. */
constexpr size_t RHS_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) double scratch[RHS_SCRATCH_BYTES / sizeof(double)];

struct bfamx_subdomain
{
  bfamx_device_t *dev;
  int physics, prec, dim, N, Np, nf;
  int32_t K;
  int state_fp32, rate_fp32;
  SubDev h;        /* host image of the descriptor */
  SubDev *d_sub;   /* device copy */
  Task *d_tasks_node; /* elastic fused: tasks of the node-parallel kernel */
  int ntasks_node;
  Task *d_tasks;
  int *d_ntasks;
  int ntasks;
  Clock *d_clock;
  double *d_partial; /* 4K reduction partials */
  double *d_rhs;     /* lazily allocated for bfamx_rhs on fused precisions */
  std::vector<void *> allocs;
  size_t bytes;
  int fields_set;
  /* host record of the clock: t0 and dt of the last set_fields and the
   * stages launched since (glued batch members must agree on all three) */
  double t0_set, dt_set;
  long long stages_run;
  bfamx_lsrk_t lsrk;
  bfamx_glue_t *glue; /* set while faces of this subdomain are glued */
};

struct bfamx_glue
{
  bfamx_device_t *dev;
  std::vector<bfamx_subdomain_t *> subs;
  int prec;
  MortarGroup *d_groups;
  int ngroups;
  void *d_ops;
  void *d_buf;
  std::vector<void *> allocs;
  size_t bytes;
  /* original face tables of the glued faces, restored by bfamx_glue_free */
  struct saved_face
  {
    int sub;
    int32_t idx, elem;
    uint8_t code;
  };
  std::vector<saved_face> saved;
};

struct bfamx_batch
{
  bfamx_device_t *dev;
  std::vector<bfamx_subdomain_t *> subs;
  SubDev *d_subs;
  Task *d_tasks;
  int *d_ntasks;
  int ntasks;
  int physics, prec, N; /* N = -1: mixed orders */
  size_t bytes;
  hipGraph_t graph;
  hipGraphExec_t exec;
  int graph_steps;
  bfamx_glue_t *glue;
  /* mixed orders: tasks sorted by order; [off, off+cnt) of order N */
  int grouped;
  int order_mask;
  int grp_off[9], grp_cnt[9];
  /* the same for the node-parallel kernel's tasks (elastic fused) */
  Task *d_tasks_node;
  int ntasks_node;
  int grpn_off[9], grpn_cnt[9];
};

#define HIPCHK(call)                                                           \
  do                                                                           \
  {                                                                            \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return bfamx_fail(BFAMX_EHIP, "%s failed: %s", #call,                   \
                        hipGetErrorString(e_));                                \
  } while (0)

/* ---------------------------------------------------------------------- */
/* device                                                                  */
/* ---------------------------------------------------------------------- */

extern "C" int bfamx_device_create(int hip_device, size_t hbm_budget_bytes,
                                   bfamx_device_t **out)
{
  if (!out)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_device_create: NULL out");
  *out = NULL;
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (hip_device < 0 || hip_device >= ndev)
    return bfamx_fail(BFAMX_EINVAL, "no HIP device %d (%d present)",
                      hip_device, ndev);
  HIPCHK(hipSetDevice(hip_device));
  size_t free_b = 0, total_b = 0;
  HIPCHK(hipMemGetInfo(&free_b, &total_b));
  bfamx_device_t *d = new bfamx_device_t();
  d->hip_dev = hip_device;
  d->budget = hbm_budget_bytes ? hbm_budget_bytes : (size_t)(0.9 * total_b);
  d->used = 0;
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, hip_device));
  d->num_cu = prop.multiProcessorCount;
  int nx = 1;
  if (hipDeviceGetAttribute(&nx, hipDeviceAttributeNumberOfXccs, hip_device) !=
          hipSuccess ||
      nx < 1)
    nx = 1;
  d->num_xcd = nx;
  /* block order: one round-robin front over the elements, each XCD taking
   * whole 2x2x2 Morton cubes (48^3 N=7 stage: FP32 1.48 ms, against 1.50
   * for the plain round-robin order and 1.55 for contiguous XCD slices;
   * FP64 2.79, 2.87 and 2.99 ms) */
  d->xcd_remap = 2;
  memset(d->mixed_choice, -1, sizeof(d->mixed_choice));
  d->persistent = 0;
  memset(d->line_bpc, 0, sizeof(d->line_bpc));
  memset(d->kchoice, -1, sizeof(d->kchoice));
  memset(d->occ, 0, sizeof(d->occ));
  memset(d->occ_dyn, 0xff, sizeof(d->occ_dyn));
  /* this device's copy of the no-allocation staging buffer, looked up once
   * with the device current */
  d->scratch = NULL;
  {
    void *ptr = NULL;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
    {
      delete d;
      return bfamx_fail(BFAMX_EHIP, "bfamx_device_create: scratch lookup failed: %s",
                        hipGetErrorString(e));
    }
    d->scratch = (double *)ptr;
  }
  /* default priority only: streams of different priorities have hung on
   * this hardware */
  HIPCHK(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&d->ev0));
  HIPCHK(hipEventCreate(&d->ev1));
  HIPCHK(hipEventCreateWithFlags(&d->fork_ev, hipEventDisableTiming));
  for (int i = 0; i < 8; ++i)
  {
    HIPCHK(hipStreamCreateWithFlags(&d->side[i], hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&d->join_ev[i], hipEventDisableTiming));
  }
  *out = d;
  return BFAMX_OK;
}

extern "C" int bfamx_device_destroy(bfamx_device_t *d)
{
  if (!d)
    return BFAMX_OK;
  (void)hipSetDevice(d->hip_dev);
  (void)hipStreamSynchronize(d->stream);
  (void)hipEventDestroy(d->ev0);
  (void)hipEventDestroy(d->ev1);
  (void)hipEventDestroy(d->fork_ev);
  for (int i = 0; i < 8; ++i)
  {
    (void)hipStreamSynchronize(d->side[i]);
    (void)hipStreamDestroy(d->side[i]);
    (void)hipEventDestroy(d->join_ev[i]);
  }
  (void)hipStreamDestroy(d->stream);
  delete d;
  return BFAMX_OK;
}

extern "C" int bfamx_device_memory(const bfamx_device_t *d, size_t *used,
                                   size_t *budget)
{
  if (!d)
    return bfamx_fail(BFAMX_EINVAL, "NULL device");
  if (used)
    *used = d->used;
  if (budget)
    *budget = d->budget;
  return BFAMX_OK;
}

extern "C" int bfamx_device_synchronize(bfamx_device_t *d)
{
  if (!d)
    return bfamx_fail(BFAMX_EINVAL, "NULL device");
  HIPCHK(hipStreamSynchronize(d->stream));
  return BFAMX_OK;
}

extern "C" void *bfamx_device_stream(bfamx_device_t *d)
{
  return d ? (void *)d->stream : NULL;
}

extern "C" int bfamx_device_info(const bfamx_device_t *d, int *num_cu,
                                 int *num_xcd)
{
  if (!d)
    return bfamx_fail(BFAMX_EINVAL, "NULL device");
  if (num_cu)
    *num_cu = d->num_cu;
  if (num_xcd)
    *num_xcd = d->num_xcd;
  return BFAMX_OK;
}

extern "C" int bfamx_set_tuning(bfamx_device_t *d, int xcd_remap)
{
  if (!d)
    return bfamx_fail(BFAMX_EINVAL, "NULL device");
  if (xcd_remap < 0 || xcd_remap > 2)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_set_tuning: mode 0, 1 or 2");
  d->xcd_remap = xcd_remap;
  return BFAMX_OK;
}

extern "C" int bfamx_timer_start(bfamx_device_t *d)
{
  if (!d)
    return bfamx_fail(BFAMX_EINVAL, "NULL device");
  HIPCHK(hipEventRecord(d->ev0, d->stream));
  return BFAMX_OK;
}

extern "C" int bfamx_timer_stop(bfamx_device_t *d, float *ms)
{
  if (!d || !ms)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_timer_stop: bad arguments");
  HIPCHK(hipEventRecord(d->ev1, d->stream));
  HIPCHK(hipEventSynchronize(d->ev1));
  HIPCHK(hipEventElapsedTime(ms, d->ev0, d->ev1));
  return BFAMX_OK;
}

/* budgeted allocation (bfamo_device_malloc, bfam.c:8561-8589) */
static int dev_alloc(bfamx_device_t *d, std::vector<void *> &list,
                     size_t *acc, size_t bytes, void **p)
{
  *p = NULL;
  if (bytes == 0)
    bytes = 16;
  bytes = (bytes + 255) & ~(size_t)255;
  if (d->used + bytes > d->budget)
    return bfamx_fail(BFAMX_EBUDGET,
                      "HBM budget exceeded: %zu used + %zu requested > %zu",
                      d->used, bytes, d->budget);
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess)
    return bfamx_fail(BFAMX_EHIP, "hipMalloc(%zu) failed: %s", bytes,
                      hipGetErrorString(e));
  d->used += bytes;
  *acc += bytes;
  list.push_back(*p);
  return BFAMX_OK;
}

static void dev_free_all(bfamx_device_t *d, std::vector<void *> &list,
                         size_t bytes)
{
  for (void *p : list)
    (void)hipFree(p);
  list.clear();
  d->used -= bytes;
}

/* ---------------------------------------------------------------------- */
/* subdomain                                                              */
/* ---------------------------------------------------------------------- */

static int nfields_of(int physics) { return physics == BFAMX_ACOUSTIC2D ? 3 : 9; }

/* elements per task: the fused elastic stage runs one wavefront per task
 * (line formulation), the unfused path one node per thread */
static int epb_of(int physics, int prec, int N)
{
  if (physics == BFAMX_ACOUSTIC2D)
    return acoustic_epb(N);
  return prec == BFAMX_FP64_PLAIN ? elastic_epb(N) : elastic_line_epw(N);
}

static int order_supported(int physics, int N)
{
  return physics == BFAMX_ACOUSTIC2D ? (N >= 1 && N <= 8) : (N >= 3 && N <= 7);
}

/* FP32 and FP64 keep no dq between stages when the table allows it:
 * A[0] = 0 (the first stage reads no dq) and B[0..S-2] nonzero (so that
 * dq_{s-1} = (q_s - q_{s-1}) / B_{s-1}) */
static int rebuilds_dq(int prec, const bfamx_lsrk_t &t)
{
  if (prec != BFAMX_FP32 && prec != BFAMX_FP64 && prec != BFAMX_FP32_BF16X3)
    return 0;
  if (t.A[0] != 0.0)
    return 0;
  for (int i = 0; i + 1 < t.n_stages; ++i)
    if (t.B[i] == 0.0)
      return 0;
  return 1;
}

static size_t subdomain_bytes(int dim, int N, int32_t K, int ph, int prec, int dq_stored);

extern "C" size_t bfamx_subdomain_bytes(int dim, int N, int32_t K,
                                        bfamx_physics_t ph,
                                        bfamx_precision_t prec)
{
  bfamx_lsrk_t t;
  bfamx_lsrk_init(&t, BFAMX_LSRK_KC54);
  return bfamx_subdomain_bytes_lsrk(dim, N, K, ph, prec, &t);
}

extern "C" size_t bfamx_subdomain_bytes_lsrk(int dim, int N, int32_t K,
                                             bfamx_physics_t ph,
                                             bfamx_precision_t prec,
                                             const bfamx_lsrk_t *lsrk)
{
  bfamx_lsrk_t t;
  if (lsrk)
    t = *lsrk;
  else
    bfamx_lsrk_init(&t, BFAMX_LSRK_KC54);
  const int dq_stored = (prec == BFAMX_FP32 || prec == BFAMX_FP64 || prec == BFAMX_FP32_BF16X3) &&
                        !rebuilds_dq(prec, t);
  return subdomain_bytes(dim, N, K, ph, prec, dq_stored);
}

/* dq_stored: an FP32 or FP64 subdomain whose table needs a stored dq */
static size_t subdomain_bytes(int dim, int N, int32_t K, int ph, int prec, int dq_stored)
{
  const size_t n1 = N + 1;
  const size_t Np = dim == 2 ? n1 * n1 : n1 * n1 * n1;
  const size_t nf = nfields_of(ph);
  const size_t ndof = (size_t)K * nf * Np;
  const size_t rnd = 255;
  size_t b = 0;
  auto add = [&](size_t x) { b += ((x ? x : 16) + rnd) & ~rnd; };
  /* state and dq buffers carry 16 bytes past their values */
  auto addq = [&](size_t x) { add(x + 16); };
  switch (prec)
  {
  /* FP64 and FP32 keep no dq unless the table needs one (rebuilds_dq) */
  case BFAMX_FP64:
    addq(8 * ndof), addq(8 * ndof);
    if (dq_stored)
      addq(8 * ndof);
    break;
  case BFAMX_FP64_PLAIN: addq(8 * ndof), add(8 * ndof), addq(8 * ndof); break;
  case BFAMX_FP32_RHS: addq(8 * ndof), addq(8 * ndof), addq(4 * ndof); break;
  default:
    addq(4 * ndof), addq(4 * ndof);
    if (dq_stored)
      addq(4 * ndof);
    break;
  }
  const size_t nfaces = 2 * dim;
  add(4 * K * nfaces), add(K * nfaces);
  add(8 * K * GEO_STRIDE), add(4 * K * GEO_STRIDE);
  add(8 * K * MAT_STRIDE), add(4 * K * MAT_STRIDE);
  add(8 * K * nfaces * FMAT_STRIDE), add(4 * K * nfaces * FMAT_STRIDE);
  add(8 * n1 * n1), add(4 * n1 * n1), add(8 * n1), add(8 * n1);
  add(sizeof(Clock)), add(sizeof(SubDev));
  /* the task lists of both stage kernels, the task count, the partial
   * sums of the reductions: as bfamx_subdomain_upload allocates them */
  const int epb = epb_of(ph, prec, N);
  add(sizeof(Task) * ((K + epb - 1) / epb));
  const int epn = elastic_epb(N);
  const int has_node = ph == BFAMX_ELASTIC3D && prec != BFAMX_FP64_PLAIN;
  add(sizeof(Task) * (has_node ? (K + epn - 1) / epn : 1));
  add(sizeof(int));
  add(32 * (size_t)K);
  return b;
}

static int push_sub(bfamx_subdomain_t *s)
{
  HIPCHK(hipMemcpy(s->d_sub, &s->h, sizeof(SubDev), hipMemcpyHostToDevice));
  return BFAMX_OK;
}

extern "C" int bfamx_subdomain_upload(bfamx_device_t *dev,
                                      const bfamx_dgx_t *sub,
                                      const bfamx_physics_desc_t *phys,
                                      const bfamx_lsrk_t *lsrk,
                                      bfamx_subdomain_t **out)
{
  if (!dev || !sub || !phys || !out)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_subdomain_upload: NULL argument");
  *out = NULL;
  const int dim = sub->dim, N = sub->N;
  const int physics = phys->physics, prec = phys->precision;
  if (physics != BFAMX_ACOUSTIC2D && physics != BFAMX_ELASTIC3D)
    return bfamx_fail(BFAMX_EINVAL, "unknown physics %d", physics);
  if (dim != (physics == BFAMX_ACOUSTIC2D ? 2 : 3))
    return bfamx_fail(BFAMX_EINVAL, "dim %d does not match physics", dim);
  if (!order_supported(physics, N))
    return bfamx_fail(BFAMX_EUNSUPPORTED, "order N=%d not built for physics %d",
                      N, physics);
  if (prec < BFAMX_FP64 || prec > BFAMX_FP32_BF16X3)
    return bfamx_fail(BFAMX_EINVAL, "unknown precision %d", prec);
  if (prec == BFAMX_FP32_BF16X3 && physics != BFAMX_ELASTIC3D)
    return bfamx_fail(BFAMX_EUNSUPPORTED, "split-BF16x3 derivatives: elastic only");
  const int n1 = N + 1;
  const int Np = dim == 2 ? n1 * n1 : n1 * n1 * n1;
  if (sub->Np != Np || sub->K <= 0 || !sub->r || !sub->w || !sub->Dr ||
      !sub->vmapM || !sub->vmapP || !phys->elem_h || !phys->rho || !phys->cp ||
      (physics == BFAMX_ELASTIC3D && !phys->cs))
    return bfamx_fail(BFAMX_EINVAL, "bfamx_subdomain_upload: incomplete input");
  const int32_t K = sub->K;
  const int nf = nfields_of(physics), nfaces = 2 * dim;
  if ((double)K * nf * Np > 4.0e9)
    return bfamx_fail(BFAMX_EINVAL, "subdomain too large for 32-bit tasks");

  /* face connectivity from vmapP */
  std::vector<int32_t> nbr_e((size_t)K * nfaces);
  std::vector<uint8_t> nbr_c((size_t)K * nfaces);
  int rc = bfamx_face_connectivity(dim, N, K, sub->vmapM, sub->vmapP,
                                   nbr_e.data(), nbr_c.data());
  if (rc)
    return rc;
  /* The kernels take each element's geometry as axis-aligned edge lengths,
   * so a face must meet the opposite face of its neighbour (f ^ 1) node for
   * node, or itself (a boundary or glued face).  A connectivity with
   * rotated trees (face pairs of another axis, or a permuted node order)
   * would upload and give a wrong derivative: refuse it. */
  for (int32_t e = 0; e < K; ++e)
    for (int f = 0; f < nfaces; ++f)
    {
      const size_t i = (size_t)e * nfaces + f;
      const int f2 = nbr_c[i] & 7, perm = nbr_c[i] >> 3;
      const int self = nbr_e[i] == e && f2 == f;
      if (!self && (f2 != (f ^ 1) || perm != 0))
        return bfamx_fail(BFAMX_EUNSUPPORTED,
                          "element %d face %d meets face %d of element %d with node "
                          "permutation %d: only axis-aligned connectivity (opposite "
                          "faces, identity permutation) is supported",
                          e, f, f2, nbr_e[i], perm);
    }

  /* geometry and material tables */
  std::vector<double> geo((size_t)K * GEO_STRIDE, 0.0), mat((size_t)K * MAT_STRIDE, 0.0);
  const double wend = sub->w[0];
  for (int32_t e = 0; e < K; ++e)
  {
    double *G = &geo[(size_t)e * GEO_STRIDE];
    for (int d = 0; d < dim; ++d)
    {
      const double h = phys->elem_h[(size_t)e * dim + d];
      if (!(h > 0))
        return bfamx_fail(BFAMX_EINVAL, "element %d has edge length %g", e, h);
      G[GEO_GX + d] = 2.0 / h;
      G[GEO_LX + d] = (2.0 / h) / wend;
      G[GEO_HX + d] = h;
      G[GEO_X0 + d] = phys->elem_x0 ? phys->elem_x0[(size_t)e * dim + d] : 0.0;
    }
    double *M = &mat[(size_t)e * MAT_STRIDE];
    const double rho = phys->rho[e], cp = phys->cp[e];
    const double cs = physics == BFAMX_ELASTIC3D ? phys->cs[e] : 0.0;
    /* cs = 0: a fluid element (acoustic medium in the velocity-stress
     * form); the exact Riemann flux then couples it to solid neighbours
     * (no shear traction, tangential slip).  The bulk modulus
     * rho (cp^2 - 4/3 cs^2) must be positive */
    if (!(rho > 0) || !(cp > 0) || !(cs >= 0) || !(3 * cp * cp > 4 * cs * cs))
      return bfamx_fail(BFAMX_EINVAL, "element %d has invalid material", e);
    M[MAT_RHOI] = 1.0 / rho;
    M[MAT_LAM] = rho * (cp * cp - 2 * cs * cs);
    M[MAT_MU] = rho * cs * cs;
    M[MAT_ZP] = rho * cp;
    M[MAT_ZS] = rho * cs;
  }
  /* plus-side impedances and inverse sums per element face */
  std::vector<double> fmat((size_t)K * nfaces * FMAT_STRIDE, 0.0);
  for (int32_t e = 0; e < K; ++e)
    for (int f = 0; f < nfaces; ++f)
    {
      const int32_t ep = nbr_e[(size_t)e * nfaces + f];
      const double *Mm = &mat[(size_t)e * MAT_STRIDE];
      const double *Mp = &mat[(size_t)ep * MAT_STRIDE];
      double *F = &fmat[((size_t)e * nfaces + f) * FMAT_STRIDE];
      F[FMAT_ZP] = Mp[MAT_ZP];
      F[FMAT_ZS] = Mp[MAT_ZS];
      F[FMAT_IZP] = 1.0 / (Mm[MAT_ZP] + Mp[MAT_ZP]);
      const double zs = Mm[MAT_ZS] + Mp[MAT_ZS];
      F[FMAT_IZS] = zs > 0 ? 1.0 / zs : 0.0;
    }
  std::vector<float> fmatf(fmat.begin(), fmat.end());
  std::vector<float> geof(geo.begin(), geo.end()), matf(mat.begin(), mat.end());
  std::vector<float> Drf(sub->Dr, sub->Dr + n1 * n1);

  bfamx_subdomain_t *s = new bfamx_subdomain_t();
  s->dev = dev;
  s->physics = physics;
  s->prec = prec;
  s->dim = dim;
  s->N = N;
  s->Np = Np;
  s->nf = nf;
  s->K = K;
  s->state_fp32 = prec == BFAMX_FP32 || prec == BFAMX_FP32_BF16X3;
  s->rate_fp32 = s->state_fp32 || prec == BFAMX_FP32_RHS;
  s->bytes = 0;
  s->d_rhs = NULL;
  s->fields_set = 0;
  if (lsrk)
    s->lsrk = *lsrk;
  else
    bfamx_lsrk_init(&s->lsrk, BFAMX_LSRK_KC54);
  if (s->lsrk.n_stages < 1 || s->lsrk.n_stages > 8)
  {
    delete s;
    return bfamx_fail(BFAMX_EINVAL, "bad LSRK table");
  }

  const size_t ndof = (size_t)K * nf * Np;
  const size_t ssz = s->state_fp32 ? 4 : 8, rsz = s->rate_fp32 ? 4 : 8;
  SubDev &h = s->h;
  memset(&h, 0, sizeof(h));
  h.physics = physics;
  h.N = N;
  h.K = K;
  h.nf = nf;
  h.nstages = s->lsrk.n_stages;
  h.src_elem = -1;
  h.mms = 0;
  h.wi0 = 1.0 / sub->w[0];
  for (int i = 0; i < 8; ++i)
  {
    h.A[i] = s->lsrk.A[i];
    h.B[i] = s->lsrk.B[i];
  }
  for (int i = 0; i < 9; ++i)
    h.C[i] = s->lsrk.C[i];
  /* FP32 and FP64 keep no dq between stages: stage s >= 1 rebuilds it from
   * the state it reads and the one before it, still in the other ping-pong
   * buffer.  Per node a middle stage then moves three state-sized values
   * (q in, q_old in, q out) instead of four (q in, dq in, q out, dq out) */
  h.dq_rec = rebuilds_dq(prec, s->lsrk);
  for (int i = 0; i < 8; ++i)
    h.AB[i] = !h.dq_rec ? h.A[i] : i > 0 && i < h.nstages ? h.A[i] / h.B[i - 1] : 0.0;

  const int epb = epb_of(physics, prec, N);
  s->ntasks = (K + epb - 1) / epb;
  std::vector<Task> tasks(s->ntasks);
  for (int t = 0; t < s->ntasks; ++t)
  {
    tasks[t].sub = 0;
    tasks[t].e0 = t * epb;
    tasks[t].ne = (t * epb + epb <= K) ? epb : K - t * epb;
    tasks[t].N = N;
  }
  /* the fused elastic stage has two kernels; the node-parallel one takes
   * elastic_epb(N) elements per task */
  const int has_node = physics == BFAMX_ELASTIC3D && prec != BFAMX_FP64_PLAIN;
  const int epn = elastic_epb(N);
  s->ntasks_node = has_node ? (K + epn - 1) / epn : 0;
  std::vector<Task> tasks_node(s->ntasks_node > 0 ? s->ntasks_node : 1);
  for (int t = 0; t < s->ntasks_node; ++t)
  {
    tasks_node[t].sub = 0;
    tasks_node[t].e0 = t * epn;
    tasks_node[t].ne = (t * epn + epn <= K) ? epn : K - t * epn;
    tasks_node[t].N = N;
  }

  void *p;
#define ALLOC(ptr, nb_)                                                        \
  do                                                                           \
  {                                                                            \
    rc = dev_alloc(dev, s->allocs, &s->bytes, (nb_), &p);                      \
    if (rc)                                                                    \
      goto fail;                                                               \
    ptr = (decltype(ptr))p;                                                    \
  } while (0)
  /* the states and dq carry 16 bytes past their values: the node-parallel
   * stage copies element blocks into LDS in 16-byte pieces from the
   * 16-byte boundaries around each block, which at odd Np reach up to
   * 12 bytes past the last element */
  ALLOC(h.q[0], ssz * ndof + 16);
  if (prec == BFAMX_FP64_PLAIN)
  {
    h.q[1] = h.q[0];
    ALLOC(h.rhs, 8 * ndof);
  }
  else
    ALLOC(h.q[1], ssz * ndof + 16);
  /* no dq buffer where stages rebuild it from the states (h.dq stays
   * NULL: no kernel reads or writes it then) */
  if (!h.dq_rec)
    ALLOC(h.dq, rsz * ndof + 16);
  {
    int32_t *ne;
    uint8_t *nc;
    double *g, *m, *D, *r, *w, *fm;
    float *gf, *mf, *Df, *fmf;
    ALLOC(ne, 4 * nbr_e.size());
    ALLOC(nc, nbr_c.size());
    ALLOC(g, 8 * geo.size());
    ALLOC(gf, 4 * geof.size());
    ALLOC(m, 8 * mat.size());
    ALLOC(mf, 4 * matf.size());
    ALLOC(fm, 8 * fmat.size());
    ALLOC(fmf, 4 * fmatf.size());
    ALLOC(D, 8 * n1 * n1);
    ALLOC(Df, 4 * n1 * n1);
    ALLOC(r, 8 * n1);
    ALLOC(w, 8 * n1);
    ALLOC(s->d_clock, sizeof(Clock));
    ALLOC(s->d_sub, sizeof(SubDev));
    ALLOC(s->d_tasks, sizeof(Task) * tasks.size());
    ALLOC(s->d_tasks_node, sizeof(Task) * tasks_node.size());
    ALLOC(s->d_ntasks, sizeof(int));
    ALLOC(s->d_partial, 32 * (size_t)K);
    hipError_t he = hipSuccess;
#define CP(dst, src, bytes)                                                    \
  if (he == hipSuccess)                                                        \
  he = hipMemcpy((void *)(dst), (const void *)(src), (bytes), hipMemcpyHostToDevice)
    CP(ne, nbr_e.data(), 4 * nbr_e.size());
    CP(nc, nbr_c.data(), nbr_c.size());
    CP(g, geo.data(), 8 * geo.size());
    CP(gf, geof.data(), 4 * geof.size());
    CP(m, mat.data(), 8 * mat.size());
    CP(mf, matf.data(), 4 * matf.size());
    CP(fm, fmat.data(), 8 * fmat.size());
    CP(fmf, fmatf.data(), 4 * fmatf.size());
    CP(D, sub->Dr, 8 * n1 * n1);
    CP(Df, Drf.data(), 4 * n1 * n1);
    CP(r, sub->r, 8 * n1);
    CP(w, sub->w, 8 * n1);
    CP(s->d_tasks, tasks.data(), sizeof(Task) * tasks.size());
    CP(s->d_tasks_node, tasks_node.data(), sizeof(Task) * tasks_node.size());
    CP(s->d_ntasks, &s->ntasks, sizeof(int));
#undef CP
    if (he != hipSuccess)
    {
      rc = bfamx_fail(BFAMX_EHIP, "upload failed: %s", hipGetErrorString(he));
      goto fail;
    }
    h.nbr_elem = ne;
    h.nbr_code = nc;
    h.geo = g;
    h.geo_f = gf;
    h.mat = m;
    h.mat_f = mf;
    h.fmat = fm;
    h.fmat_f = fmf;
    h.Dr = D;
    h.Dr_f = Df;
    h.r = r;
    h.w = w;
    h.clock = s->d_clock;
  }
#undef ALLOC
  {
    Clock c;
    memset(&c, 0, sizeof(c));
    if (hipMemcpy(s->d_clock, &c, sizeof(c), hipMemcpyHostToDevice) != hipSuccess)
    {
      rc = bfamx_fail(BFAMX_EHIP, "clock init failed");
      goto fail;
    }
  }
  rc = push_sub(s);
  if (rc)
    goto fail;
  *out = s;
  return BFAMX_OK;
fail:
  dev_free_all(dev, s->allocs, s->bytes);
  delete s;
  return rc;
}

extern "C" int bfamx_subdomain_free(bfamx_subdomain_t *s)
{
  if (!s)
    return BFAMX_OK;
  if (s->glue)
    return bfamx_fail(BFAMX_ESTATE, "bfamx_subdomain_free: free its glue first");
  (void)hipStreamSynchronize(s->dev->stream);
  dev_free_all(s->dev, s->allocs, s->bytes);
  delete s;
  return BFAMX_OK;
}

extern "C" int bfamx_subdomain_info(const bfamx_subdomain_t *s, int *nfields,
                                    size_t *device_bytes)
{
  if (!s)
    return bfamx_fail(BFAMX_EINVAL, "NULL subdomain");
  if (nfields)
    *nfields = s->nf;
  if (device_bytes)
    *device_bytes = s->bytes;
  return BFAMX_OK;
}

static double host_ricker(double f0, double t0, double t)
{
  const double x = 3.14159265358979323846 * f0 * (t - t0);
  const double a = x * x;
  return (1 - 2 * a) * exp(-a);
}

/* recompute the time scalars of the stage about to run (host side) */
static int refresh_clock(bfamx_subdomain_t *s)
{
  Clock c;
  HIPCHK(hipStreamSynchronize(s->dev->stream));
  HIPCHK(hipMemcpy(&c, s->d_clock, sizeof(c), hipMemcpyDeviceToHost));
  const double ts = c.t + s->h.C[(int)(c.stage % s->h.nstages)] * c.dt;
  if (s->h.src_elem >= 0)
    c.g = host_ricker(s->h.src_f0, s->h.src_t0, ts);
  if (s->h.mms)
  {
    c.sw = sin(s->h.mms_omega * ts);
    c.cw = cos(s->h.mms_omega * ts);
  }
  HIPCHK(hipMemcpy(s->d_clock, &c, sizeof(c), hipMemcpyHostToDevice));
  return BFAMX_OK;
}

extern "C" int bfamx_set_source(bfamx_subdomain_t *s, int32_t elem,
                                const double *weights, const double *amp,
                                double f0, double t0)
{
  if (!s || !weights || !amp || elem < 0 || elem >= s->K)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_set_source: bad arguments");
  int rc;
  if (!s->h.src_wt)
  {
    void *p;
    rc = dev_alloc(s->dev, s->allocs, &s->bytes, 8 * s->Np, &p);
    if (rc)
      return rc;
    s->h.src_wt = (const double *)p;
    rc = dev_alloc(s->dev, s->allocs, &s->bytes, 4 * s->Np, &p);
    if (rc)
      return rc;
    s->h.src_wt_f = (const float *)p;
  }
  std::vector<float> wf(weights, weights + s->Np);
  HIPCHK(hipMemcpy((void *)s->h.src_wt, weights, 8 * s->Np, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((void *)s->h.src_wt_f, wf.data(), 4 * s->Np, hipMemcpyHostToDevice));
  s->h.src_elem = elem;
  for (int f = 0; f < 9; ++f)
    s->h.src_amp[f] = f < s->nf ? amp[f] : 0.0;
  s->h.src_f0 = f0;
  s->h.src_t0 = t0;
  rc = push_sub(s);
  if (rc)
    return rc;
  return refresh_clock(s);
}

extern "C" int bfamx_set_forcing_mms(bfamx_subdomain_t *s, double k,
                                     double omega, const double *F)
{
  if (!s || !F || s->physics != BFAMX_ELASTIC3D)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_set_forcing_mms: elastic only");
  const int n1 = s->N + 1;
  if (!s->h.mms_tab)
  {
    /* sin/cos(k x) at the element's node coordinates, per axis (static in
     * time, so computed once here instead of in every stage) */
    std::vector<double> geo((size_t)s->K * GEO_STRIDE), r(n1);
    HIPCHK(hipMemcpy(geo.data(), s->h.geo, 8 * geo.size(), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(r.data(), s->h.r, 8 * n1, hipMemcpyDeviceToHost));
    std::vector<double> tab((size_t)s->K * 6 * n1);
    for (int32_t e = 0; e < s->K; ++e)
      for (int ax = 0; ax < 3; ++ax)
        for (int i = 0; i < n1; ++i)
        {
          const double *G = &geo[(size_t)e * GEO_STRIDE];
          const double x = G[GEO_X0 + ax] + 0.5 * (r[i] + 1) * G[GEO_HX + ax];
          tab[((size_t)e * 3 + ax) * 2 * n1 + i] = sin(k * x);
          tab[((size_t)e * 3 + ax) * 2 * n1 + n1 + i] = cos(k * x);
        }
    void *p;
    int rc = dev_alloc(s->dev, s->allocs, &s->bytes, 8 * tab.size(), &p);
    if (rc)
      return rc;
    HIPCHK(hipMemcpy(p, tab.data(), 8 * tab.size(), hipMemcpyHostToDevice));
    s->h.mms_tab = (const double *)p;
    std::vector<float> tabf(tab.begin(), tab.end());
    rc = dev_alloc(s->dev, s->allocs, &s->bytes, 4 * tabf.size(), &p);
    if (rc)
      return rc;
    HIPCHK(hipMemcpy(p, tabf.data(), 4 * tabf.size(), hipMemcpyHostToDevice));
    s->h.mms_tab_f = (const float *)p;
  }
  s->h.mms = 1;
  s->h.mms_k = k;
  s->h.mms_omega = omega;
  for (int f = 0; f < 9; ++f)
    s->h.mms_F[f] = F[f];
  int rc = push_sub(s);
  if (rc)
    return rc;
  return refresh_clock(s);
}

extern "C" int bfamx_set_fields(bfamx_subdomain_t *s,
                                const double *const *fields, double t0,
                                double dt)
{
  if (!s)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_set_fields: NULL argument");
  if (fields)
    for (int f = 0; f < s->nf; ++f)
      if (!fields[f])
        return bfamx_fail(BFAMX_EINVAL, "bfamx_set_fields: field %d is NULL", f);
  const size_t Np = s->Np, nf = s->nf, K = s->K;
  const size_t ndof = K * nf * Np;
  HIPCHK(hipStreamSynchronize(s->dev->stream));
  if (!fields)
    HIPCHK(hipMemset(s->h.q[0], 0, (s->state_fp32 ? 4 : 8) * ndof));
  else if (s->state_fp32)
  {
    std::vector<float> buf(ndof);
    for (size_t e = 0; e < K; ++e)
      for (size_t f = 0; f < nf; ++f)
        for (size_t n = 0; n < Np; ++n)
          buf[(e * nf + f) * Np + n] = (float)fields[f][e * Np + n];
    HIPCHK(hipMemcpy(s->h.q[0], buf.data(), 4 * ndof, hipMemcpyHostToDevice));
  }
  else
  {
    std::vector<double> buf(ndof);
    for (size_t e = 0; e < K; ++e)
      for (size_t f = 0; f < nf; ++f)
        for (size_t n = 0; n < Np; ++n)
          buf[(e * nf + f) * Np + n] = fields[f][e * Np + n];
    HIPCHK(hipMemcpy(s->h.q[0], buf.data(), 8 * ndof, hipMemcpyHostToDevice));
  }
  if (s->h.dq)
    HIPCHK(hipMemset(s->h.dq, 0, (s->rate_fp32 ? 4 : 8) * ndof));
  Clock c;
  memset(&c, 0, sizeof(c));
  c.t = t0;
  c.dt = dt;
  HIPCHK(hipMemcpy(s->d_clock, &c, sizeof(c), hipMemcpyHostToDevice));
  s->fields_set = 1;
  s->t0_set = t0;
  s->dt_set = dt;
  s->stages_run = 0;
  return refresh_clock(s);
}

static int current_parity(bfamx_subdomain_t *s, int *parity)
{
  Clock c;
  HIPCHK(hipStreamSynchronize(s->dev->stream));
  HIPCHK(hipMemcpy(&c, s->d_clock, sizeof(c), hipMemcpyDeviceToHost));
  *parity = s->prec == BFAMX_FP64_PLAIN ? 0 : (int)(c.stage & 1);
  return BFAMX_OK;
}

extern "C" int bfamx_get_time(bfamx_subdomain_t *s, double *t, int64_t *stages)
{
  if (!s)
    return bfamx_fail(BFAMX_EINVAL, "NULL subdomain");
  Clock c;
  HIPCHK(hipStreamSynchronize(s->dev->stream));
  HIPCHK(hipMemcpy(&c, s->d_clock, sizeof(c), hipMemcpyDeviceToHost));
  if (t)
    *t = c.t;
  if (stages)
    *stages = c.stage;
  return BFAMX_OK;
}

/* copy a device [K][nf][Np] array of T into bfam's per-field host layout */
template <class T>
static int download_layout(const bfamx_subdomain_t *s, const void *src,
                           double *const *fields)
{
  const size_t Np = s->Np, nf = s->nf, K = s->K;
  std::vector<T> buf(K * nf * Np);
  HIPCHK(hipMemcpy(buf.data(), src, sizeof(T) * buf.size(), hipMemcpyDeviceToHost));
  for (size_t e = 0; e < K; ++e)
    for (size_t f = 0; f < nf; ++f)
      for (size_t n = 0; n < Np; ++n)
        fields[f][e * Np + n] = (double)buf[(e * nf + f) * Np + n];
  return BFAMX_OK;
}

extern "C" int bfamx_download_fields(bfamx_subdomain_t *s,
                                     double *const *fields)
{
  if (!s || !fields)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_download_fields: NULL argument");
  if (!s->fields_set)
    return bfamx_fail(BFAMX_ESTATE, "bfamx_download_fields: no state set");
  int parity, rc = current_parity(s, &parity);
  if (rc)
    return rc;
  return s->state_fp32 ? download_layout<float>(s, s->h.q[parity], fields)
                       : download_layout<double>(s, s->h.q[parity], fields);
}

static LaunchArgs launch_args(bfamx_subdomain_t *s, int mode)
{
  LaunchArgs a;
  memset(&a, 0, sizeof(a));
  a.subs = s->d_sub;
  a.tasks = s->d_tasks;
  a.ntasks = s->ntasks;
  a.mode = mode;
  a.xcd_remap = s->dev->xcd_remap;
  a.nxcd = s->dev->num_xcd;
  return a;
}

/* grid of an elastic stage launch over ntasks tasks of order N (N < 0:
 * the order-switching kernel) */
static int grid_line(bfamx_device_t *d, int prec, int N, int ntasks);

static int grid_for(const bfamx_device_t *d, int ntasks)
{
  if (d->xcd_remap != 1 || d->num_xcd <= 1)
    return ntasks;
  return d->num_xcd * ((ntasks + d->num_xcd - 1) / d->num_xcd);
}

/* the stage kernel of an order not (yet) autotuned (0 line, 1
 * node-parallel, 2 node-parallel on a resident grid): the one measured
 * fastest on bricks of about 124 M DOF (FP32: resident grid at N = 4 and
 * 6, node-parallel at N = 3, 5 and 7; FP64: node-parallel at N = 3..5,
 * resident grid at N = 6, 7; at N = 5 the two node-parallel grids are
 * within 5 % of each other), and the node-parallel kernel for FP32_RHS and
 * split-BF16x3 */
static int default_kind(int prec, int N)
{
  switch (prec)
  {
  case PREC_FP32: return (N == 4 || N == 6) ? 2 : 1;
  case PREC_FP32_RHS: return 1;
  case PREC_FP32_BF16X3: return 1;
  case PREC_FP64: return N <= 5 ? 1 : 2;
  default: return 0;
  }
}

static int kernel_kind(const bfamx_device_t *d, int prec, int N)
{
  if (N < 3 || N > 7)
    return 0;
  const signed char k = d->kchoice[prec][N];
  return k < 0 ? default_kind(prec, N) : k;
}
static int node_kernel(const bfamx_device_t *d, int prec, int N)
{
  return kernel_kind(d, prec, N) >= 1;
}
/* the node-parallel kernel of (prec, N) runs on a resident grid */
static int node_resident(const bfamx_device_t *d, int prec, int N)
{
  const int k = kernel_kind(d, prec, N);
  return k == 2 || k == 4;
}
/* the node-parallel kernel of (prec, N) runs its other block size */
static int node_alt(const bfamx_device_t *d, int prec, int N)
{
  return kernel_kind(d, prec, N) >= 3;
}
/* occupancy-cap slot of a kernel kind: 0 line, 1 node-parallel, 2
 * node-parallel with the other block size */
static int kind_slot(int kind)
{
  return kind == 0 ? 0 : (kind <= 2 ? 1 : 2);
}
static int kernel_slot(const bfamx_device_t *d, int prec, int N)
{
  return kind_slot(kernel_kind(d, prec, N));
}

/* dynamic LDS bytes per block that cap the stage kernel of (prec, N,
 * slot: 0 line, 1 node-parallel, 2 node-parallel with its other block
 * size) at level lv (0: no cap, 1: 3/4 of the uncapped resident blocks
 * per CU, 2: 1/2): the smallest request at which the occupancy calculator
 * allows no more than that many blocks; 0 if the cap is not below the
 * uncapped count or cannot be computed */
static unsigned occ_dyn(bfamx_device_t *d, int prec, int N, int slot, int lv)
{
  if (lv <= 0 || N < 3 || N > 7)
    return 0;
  unsigned &c = d->occ_dyn[prec][N][slot][lv];
  if (c != ~0u)
    return c;
  c = 0;
  auto blocks = [&](unsigned dyn, int *nb) {
    return slot ? node_blocks_per_cu(prec, N, nb, dyn, slot == 2)
                : line_blocks_per_cu(prec, N, nb, dyn);
  };
  int b0 = 0;
  if (blocks(0, &b0) != hipSuccess || b0 < 2)
    return c;
  const int target = lv == 1 ? (3 * b0) / 4 : b0 / 2;
  if (target < 1 || target >= b0)
    return c;
  unsigned lo = 0, hi = 160 * 1024; /* blocks(hi) <= target: at most one block */
  while (hi - lo > 256)
  {
    const unsigned mid = (lo + hi) / 2;
    int nb = 0;
    if (blocks(mid, &nb) != hipSuccess)
      return c;
    (nb <= target ? hi : lo) = mid;
  }
  int nb = 0;
  if (blocks(hi, &nb) == hipSuccess && nb >= 1 && nb <= target)
    c = hi;
  return c;
}
static unsigned kernel_dyn(bfamx_device_t *d, int prec, int N, int slot)
{
  return N >= 3 && N <= 7 ? occ_dyn(d, prec, N, slot, d->occ[prec][N][slot]) : 0;
}
/* grid of a node-parallel launch over ntasks tasks: one block per task, or,
 * on a resident grid, the blocks that fit the CUs at once (each then takes
 * every grid-th task; its setup and block launch are paid once per CU slot
 * instead of once per task) */
static int grid_node(bfamx_device_t *d, int prec, int N, int ntasks)
{
  int g = grid_for(d, ntasks);
  if (!node_resident(d, prec, N))
    return g;
  int bpc = 0;
  const int slot = kernel_slot(d, prec, N);
  if (node_blocks_per_cu(prec, N, &bpc, kernel_dyn(d, prec, N, slot), slot == 2) ==
          hipSuccess &&
      bpc > 0 && (long long)bpc * d->num_cu < g)
    g = bpc * d->num_cu;
  return g;
}

/* does any order of the batch run the node-parallel kernel */
static int batch_any_node(const bfamx_batch_t *b);

static int run_fused(bfamx_subdomain_t *s, const LaunchArgs &a)
{
  hipError_t e;
  if (s->physics != BFAMX_ELASTIC3D)
    e = launch_acoustic_stage(s->prec, s->N, a, grid_for(s->dev, a.ntasks), s->dev->stream);
  else if (node_kernel(s->dev, s->prec, s->N))
  {
    LaunchArgs b = a;
    b.tasks = s->d_tasks_node;
    b.ntasks = s->ntasks_node;
    const int slot = kernel_slot(s->dev, s->prec, s->N);
    e = launch_elastic_stage(s->prec, s->N, b, grid_node(s->dev, s->prec, s->N, b.ntasks),
                             s->dev->stream, kernel_dyn(s->dev, s->prec, s->N, slot), slot == 2);
  }
  else
    e = launch_elastic_line(s->prec, s->N, a, grid_line(s->dev, s->prec, s->N, a.ntasks),
                            s->dev->stream, kernel_dyn(s->dev, s->prec, s->N, 0));
  if (e == hipSuccess && a.mode == MODE_STAGE)
    e = launch_clock_advance(s->d_sub, 1, s->dev->stream);
  if (e != hipSuccess)
    return bfamx_fail(BFAMX_EHIP, "stage launch failed: %s", hipGetErrorString(e));
  return BFAMX_OK;
}

static int run_plain_rhs(bfamx_subdomain_t *s, const LaunchArgs &a)
{
  const int grid = grid_for(s->dev, a.ntasks);
  hipError_t e = s->physics == BFAMX_ELASTIC3D
                     ? launch_elastic_plain(s->N, a, grid, s->dev->stream)
                     : launch_acoustic_plain(s->N, a, grid, s->dev->stream);
  if (e != hipSuccess)
    return bfamx_fail(BFAMX_EHIP, "plain launch failed: %s", hipGetErrorString(e));
  return BFAMX_OK;
}

static int stage_once(bfamx_subdomain_t *s)
{
  if (s->prec == BFAMX_FP64_PLAIN)
  {
    int rc = run_plain_rhs(s, launch_args(s, MODE_STAGE));
    if (rc)
      return rc;
    const size_t ndof = (size_t)s->K * s->nf * s->Np;
    const int nb = (int)((ndof + 255) / 256 < (size_t)s->dev->num_cu * 16
                             ? (ndof + 255) / 256
                             : (size_t)s->dev->num_cu * 16);
    hipError_t e = launch_plain_update(s->d_sub, ndof, s->dev->stream, nb);
    if (e == hipSuccess)
      e = launch_clock_advance(s->d_sub, 1, s->dev->stream);
    if (e != hipSuccess)
      return bfamx_fail(BFAMX_EHIP, "update launch failed: %s", hipGetErrorString(e));
    return BFAMX_OK;
  }
  return run_fused(s, launch_args(s, MODE_STAGE));
}

extern "C" int bfamx_stage(bfamx_subdomain_t *s)
{
  if (!s)
    return bfamx_fail(BFAMX_EINVAL, "NULL subdomain");
  if (!s->fields_set)
    return bfamx_fail(BFAMX_ESTATE, "bfamx_stage: no state set");
  if (s->glue)
    return bfamx_fail(BFAMX_ESTATE, "glued subdomains advance through their batch");
  int rc = stage_once(s);
  if (!rc)
    ++s->stages_run;
  return rc;
}

extern "C" int bfamx_step(bfamx_subdomain_t *s, int nsteps)
{
  if (!s || nsteps < 0)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_step: bad arguments");
  if (!s->fields_set)
    return bfamx_fail(BFAMX_ESTATE, "bfamx_step: no state set");
  if (s->glue)
    return bfamx_fail(BFAMX_ESTATE, "glued subdomains advance through their batch");
  for (int i = 0; i < nsteps; ++i)
    for (int st = 0; st < s->h.nstages; ++st)
    {
      int rc = stage_once(s);
      if (rc)
        return rc;
      ++s->stages_run;
    }
  return BFAMX_OK;
}

static int rhs_one(bfamx_subdomain_t *s, double t, double *const *rhs);

extern "C" int bfamx_rhs(bfamx_subdomain_t *s, double t, double *const *rhs)
{
  if (!s || !rhs)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_rhs: NULL argument");
  if (!s->fields_set)
    return bfamx_fail(BFAMX_ESTATE, "bfamx_rhs: no state set");
  if (s->glue)
    return bfamx_fail(BFAMX_ESTATE, "glued subdomains: use bfamx_batch_rhs");
  return rhs_one(s, t, rhs);
}

/* the MODE_RHS launch of one subdomain, writing L(q, t) as FP64 into the
 * buffer the descriptor's rhs names */
static int rhs_launch(bfamx_subdomain_t *s, double t)
{
  LaunchArgs a = launch_args(s, MODE_RHS);
  a.t_rhs = t;
  a.g_rhs = s->h.src_elem >= 0 ? host_ricker(s->h.src_f0, s->h.src_t0, t) : 0.0;
  a.sw_rhs = s->h.mms ? sin(s->h.mms_omega * t) : 0.0;
  a.cw_rhs = s->h.mms ? cos(s->h.mms_omega * t) : 0.0;
  return s->prec == BFAMX_FP64_PLAIN ? run_plain_rhs(s, a) : run_fused(s, a);
}

/*
 * bfamx_rhs without any allocation: the RHS is staged in this GPU's copy
 * of `scratch` (64 MiB) instead of the subdomain's lazily allocated FP64
 * buffer, and copied field by field into the caller's arrays with strided
 * copies.  The RHS kernels write every entry of the staging area they use
 * (the unfused volume kernel assigns before the surface kernel adds), so
 * nothing needs clearing.  Calls that overlap on one GPU share the buffer
 * and race.
 */
extern "C" int bfamx_rhs_noalloc(bfamx_subdomain_t *s, double t, double *const *rhs)
{
  if (!s || !rhs)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_rhs_noalloc: NULL argument");
  for (int f = 0; f < s->nf; ++f)
    if (!rhs[f])
      return bfamx_fail(BFAMX_EINVAL, "bfamx_rhs_noalloc: field %d is NULL", f);
  if (!s->fields_set)
    return bfamx_fail(BFAMX_ESTATE, "bfamx_rhs_noalloc: no state set");
  if (s->glue)
    return bfamx_fail(BFAMX_ESTATE, "glued subdomains: use bfamx_batch_rhs");
  const size_t Np = s->Np, nf = s->nf, K = s->K;
  const size_t ndof = K * nf * Np;
  if (8 * ndof > RHS_SCRATCH_BYTES)
    return bfamx_fail(BFAMX_EINVAL,
                      "bfamx_rhs_noalloc: %zu FP64 values exceed the %zu-byte scratch "
                      "(at most %zu values)", ndof, RHS_SCRATCH_BYTES,
                      RHS_SCRATCH_BYTES / sizeof(double));
  double *const stage = s->dev->scratch;
  double *const keep = s->h.rhs;
  /* the descriptor changes under no launch of this subdomain */
  HIPCHK(hipStreamSynchronize(s->dev->stream));
  s->h.rhs = stage;
  int rc = push_sub(s);
  if (!rc)
    rc = rhs_launch(s, t);
  hipError_t e = hipSuccess;
  if (!rc)
    e = hipStreamSynchronize(s->dev->stream);
  /* [K][nf][Np] on the device to rhs[f][K * Np] on the host */
  for (size_t f = 0; f < nf && !rc && e == hipSuccess; ++f)
    e = hipMemcpy2D(rhs[f], Np * sizeof(double), stage + f * Np, nf * Np * sizeof(double),
                    Np * sizeof(double), K, hipMemcpyDeviceToHost);
  s->h.rhs = keep;
  const int rc2 = push_sub(s);
  if (!rc && e != hipSuccess)
    rc = bfamx_fail(BFAMX_EHIP, "bfamx_rhs_noalloc: %s", hipGetErrorString(e));
  return rc ? rc : rc2;
}

/* RHS of one subdomain into host arrays; glued faces use the glue buffer as
 * the last mortar launch left it */
static int rhs_one(bfamx_subdomain_t *s, double t, double *const *rhs)
{
  const size_t ndof = (size_t)s->K * s->nf * s->Np;
  int rc;
  if (!s->h.rhs)
  {
    void *p;
    rc = dev_alloc(s->dev, s->allocs, &s->bytes, 8 * ndof, &p);
    if (rc)
      return rc;
    s->h.rhs = (double *)p;
    rc = push_sub(s);
    if (rc)
      return rc;
  }
  rc = rhs_launch(s, t);
  if (rc)
    return rc;
  HIPCHK(hipStreamSynchronize(s->dev->stream));
  return download_layout<double>(s, s->h.rhs, rhs);
}

/* Neumaier's compensated sum of n values with stride */
static double neumaier(const double *x, size_t n, size_t stride)
{
  double s = 0, c = 0;
  for (size_t i = 0; i < n; ++i)
  {
    const double v = x[i * stride];
    const double t = s + v;
    if (fabs(s) >= fabs(v))
      c += (s - t) + v;
    else
      c += (v - t) + s;
    s = t;
  }
  return s + c;
}

extern "C" int bfamx_reduce_energy(bfamx_subdomain_t *s, double *energy)
{
  if (!s || !energy)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_reduce_energy: NULL argument");
  int parity, rc = current_parity(s, &parity);
  if (rc)
    return rc;
  hipError_t e = launch_energy(s->h, s->state_fp32, parity, s->d_partial,
                               s->dev->stream);
  if (e != hipSuccess)
    return bfamx_fail(BFAMX_EHIP, "energy launch: %s", hipGetErrorString(e));
  std::vector<double> part(s->K);
  HIPCHK(hipStreamSynchronize(s->dev->stream));
  HIPCHK(hipMemcpy(part.data(), s->d_partial, 8 * part.size(), hipMemcpyDeviceToHost));
  *energy = neumaier(part.data(), part.size(), 1);
  return BFAMX_OK;
}

extern "C" int bfamx_reduce_cfl_dt(bfamx_subdomain_t *s, double cfl, double *dt)
{
  if (!s || !dt || !(cfl > 0))
    return bfamx_fail(BFAMX_EINVAL, "bfamx_reduce_cfl_dt: bad arguments");
  hipError_t e = launch_cfl(s->h, s->d_partial, s->dev->stream);
  if (e != hipSuccess)
    return bfamx_fail(BFAMX_EHIP, "cfl launch: %s", hipGetErrorString(e));
  std::vector<double> part(s->K);
  HIPCHK(hipStreamSynchronize(s->dev->stream));
  HIPCHK(hipMemcpy(part.data(), s->d_partial, 8 * part.size(), hipMemcpyDeviceToHost));
  double m = part.empty() ? 0.0 : part[0];
  for (double v : part)
    m = v < m ? v : m;
  const double n1 = s->N + 1;
  *dt = cfl * m / (n1 * n1);
  return BFAMX_OK;
}

extern "C" int bfamx_reduce_l2_diff(bfamx_subdomain_t *a, bfamx_subdomain_t *b,
                                    double *diff2, double *norm2)
{
  if (!a || !b || !diff2 || !norm2)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_reduce_l2_diff: NULL argument");
  if (a->K != b->K || a->N != b->N || a->physics != b->physics || a->dev != b->dev)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_reduce_l2_diff: different shapes");
  int pa, pb, rc = current_parity(a, &pa);
  if (rc || (rc = current_parity(b, &pb)))
    return rc;
  hipError_t e = launch_l2diff(a->h, a->state_fp32, pa, b->h, b->state_fp32, pb,
                               a->d_partial, a->dev->stream);
  if (e != hipSuccess)
    return bfamx_fail(BFAMX_EHIP, "l2diff launch: %s", hipGetErrorString(e));
  std::vector<double> part(2 * (size_t)a->K);
  HIPCHK(hipStreamSynchronize(a->dev->stream));
  HIPCHK(hipMemcpy(part.data(), a->d_partial, 8 * part.size(), hipMemcpyDeviceToHost));
  *diff2 = neumaier(part.data(), a->K, 2);
  *norm2 = neumaier(part.data() + 1, a->K, 2);
  return BFAMX_OK;
}

static int mms_partials(bfamx_subdomain_t *s, const double *A, double *out4)
{
  if (!s || !A || !s->h.mms_tab)
    return bfamx_fail(BFAMX_EINVAL, "MMS error: no forcing set");
  Clock c;
  HIPCHK(hipStreamSynchronize(s->dev->stream));
  HIPCHK(hipMemcpy(&c, s->d_clock, sizeof(c), hipMemcpyDeviceToHost));
  const int parity = s->prec == BFAMX_FP64_PLAIN ? 0 : (int)(c.stage & 1);
  /* the state is at a step boundary only when stage % nstages == 0 */
  hipError_t e = launch_mms_error(s->h, s->state_fp32, parity, A,
                                  cos(s->h.mms_omega * c.t),
                                  sin(s->h.mms_omega * c.t), s->d_partial,
                                  s->dev->stream);
  if (e != hipSuccess)
    return bfamx_fail(BFAMX_EHIP, "mms error launch: %s", hipGetErrorString(e));
  std::vector<double> part(4 * (size_t)s->K);
  HIPCHK(hipStreamSynchronize(s->dev->stream));
  HIPCHK(hipMemcpy(part.data(), s->d_partial, 8 * part.size(), hipMemcpyDeviceToHost));
  for (int k = 0; k < 4; ++k)
    out4[k] = neumaier(part.data() + k, s->K, 4);
  return BFAMX_OK;
}

extern "C" int bfamx_reduce_mms_error(bfamx_subdomain_t *s, const double *A,
                                      double *diff2, double *norm2)
{
  if (!diff2 || !norm2)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_reduce_mms_error: NULL output");
  double o[4];
  const int rc = mms_partials(s, A, o);
  if (rc)
    return rc;
  *diff2 = o[0];
  *norm2 = o[1];
  return BFAMX_OK;
}

extern "C" int bfamx_reduce_mms_energy_error(bfamx_subdomain_t *s, const double *A,
                                             double *err_energy, double *exact_energy)
{
  if (!err_energy || !exact_energy)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_reduce_mms_energy_error: NULL output");
  double o[4];
  const int rc = mms_partials(s, A, o);
  if (rc)
    return rc;
  *err_energy = o[2];
  *exact_energy = o[3];
  return BFAMX_OK;
}

/* ---------------------------------------------------------------------- */
/* batches                                                                */
/* ---------------------------------------------------------------------- */

extern "C" int bfamx_batch_create(bfamx_device_t *dev,
                                  bfamx_subdomain_t *const *subs, int nsubs,
                                  bfamx_batch_t **out)
{
  if (!dev || !subs || nsubs < 1 || !out)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_batch_create: bad arguments");
  *out = NULL;
  const int physics = subs[0]->physics, prec = subs[0]->prec;
  int N = subs[0]->N;
  for (int i = 0; i < nsubs; ++i)
  {
    const bfamx_subdomain_t *s = subs[i];
    if (!s || s->dev != dev || s->physics != physics || s->prec != prec)
      return bfamx_fail(BFAMX_EINVAL,
                        "batch members must share device, physics, precision");
    if (prec == BFAMX_FP64_PLAIN)
      return bfamx_fail(BFAMX_EUNSUPPORTED, "batches run the fused stage only");
    /* a batch step runs the same number of stages for every member */
    if (s->h.nstages != subs[0]->h.nstages)
      return bfamx_fail(BFAMX_EINVAL,
                        "batch members must use LSRK tables of one stage count "
                        "(member %d: %d stages, member 0: %d)",
                        i, s->h.nstages, subs[0]->h.nstages);
    if (s->N != N)
      N = -1;
  }
  if (N < 0 && physics != BFAMX_ELASTIC3D)
    return bfamx_fail(BFAMX_EUNSUPPORTED, "mixed orders: elastic only");
  bfamx_batch_t *b = new bfamx_batch_t();
  b->dev = dev;
  b->subs.assign(subs, subs + nsubs);
  b->physics = physics;
  b->prec = prec;
  b->N = N;
  b->bytes = 0;
  b->graph = NULL;
  b->exec = NULL;
  b->graph_steps = 0;
  b->glue = NULL;
  std::vector<SubDev> hs(nsubs);
  std::vector<Task> tasks;
  std::vector<int> per(nsubs);
  for (int i = 0; i < nsubs; ++i)
  {
    bfamx_subdomain_t *s = subs[i];
    hs[i] = s->h;
    const int epb = epb_of(physics, prec, s->N);
    int cnt = 0;
    for (int32_t e0 = 0; e0 < s->K; e0 += epb, ++cnt)
    {
      Task t;
      t.sub = i;
      t.e0 = e0;
      t.ne = e0 + epb <= s->K ? epb : s->K - e0;
      t.N = s->N;
      tasks.push_back(t);
    }
    per[i] = cnt;
  }
  std::stable_sort(tasks.begin(), tasks.end(),
                   [](const Task &x, const Task &y) { return x.N < y.N; });
  std::vector<Task> tasks_node;
  if (physics == BFAMX_ELASTIC3D)
    for (int i = 0; i < nsubs; ++i)
    {
      const int epn = elastic_epb(subs[i]->N);
      for (int32_t e0 = 0; e0 < subs[i]->K; e0 += epn)
      {
        Task t;
        t.sub = i;
        t.e0 = e0;
        t.ne = e0 + epn <= subs[i]->K ? epn : subs[i]->K - e0;
        t.N = subs[i]->N;
        tasks_node.push_back(t);
      }
    }
  std::stable_sort(tasks_node.begin(), tasks_node.end(),
                   [](const Task &x, const Task &y) { return x.N < y.N; });
  b->ntasks_node = (int)tasks_node.size();
  for (int n = 0; n < 9; ++n)
    b->grpn_off[n] = b->grpn_cnt[n] = 0;
  for (int t = 0; t < b->ntasks_node; ++t)
  {
    const int n = tasks_node[t].N;
    if (b->grpn_cnt[n]++ == 0)
      b->grpn_off[n] = t;
  }
  if (tasks_node.empty())
    tasks_node.resize(1);
  b->ntasks = (int)tasks.size();
  b->grouped = 0;
  b->order_mask = 0;
  for (int n = 0; n < 9; ++n)
    b->grp_off[n] = b->grp_cnt[n] = 0;
  for (int t = 0; t < b->ntasks; ++t)
  {
    const int n = tasks[t].N;
    if (b->grp_cnt[n]++ == 0)
      b->grp_off[n] = t;
    b->order_mask |= 1 << n;
  }
  if (N < 0)
  {
    /* until autotuned: one launch per order, which measured faster than
     * the order-switching launch in every precision, and runs each order
     * with the kernel a single-subdomain launch of it uses */
    const signed char c = dev->mixed_choice[prec][b->order_mask & 255];
    b->grouped = c != 0 || batch_any_node(b);
  }
  std::vector<void *> allocs;
  void *p;
  int rc = dev_alloc(dev, allocs, &b->bytes, sizeof(SubDev) * nsubs, &p);
  if (!rc)
  {
    b->d_subs = (SubDev *)p;
    rc = dev_alloc(dev, allocs, &b->bytes, sizeof(Task) * tasks.size(), &p);
  }
  if (!rc)
  {
    b->d_tasks = (Task *)p;
    rc = dev_alloc(dev, allocs, &b->bytes, sizeof(Task) * tasks_node.size(), &p);
  }
  if (!rc)
  {
    b->d_tasks_node = (Task *)p;
    rc = dev_alloc(dev, allocs, &b->bytes, sizeof(int) * nsubs, &p);
  }
  if (rc)
  {
    dev_free_all(dev, allocs, b->bytes);
    delete b;
    return rc;
  }
  b->d_ntasks = (int *)p;
  if (hipMemcpy(b->d_subs, hs.data(), sizeof(SubDev) * nsubs, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_tasks, tasks.data(), sizeof(Task) * tasks.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_tasks_node, tasks_node.data(), sizeof(Task) * tasks_node.size(),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_ntasks, per.data(), sizeof(int) * nsubs, hipMemcpyHostToDevice) != hipSuccess)
  {
    dev_free_all(dev, allocs, b->bytes);
    delete b;
    return bfamx_fail(BFAMX_EHIP, "batch upload failed");
  }
  *out = b;
  return BFAMX_OK;
}

extern "C" int bfamx_batch_free(bfamx_batch_t *b)
{
  if (!b)
    return BFAMX_OK;
  (void)hipStreamSynchronize(b->dev->stream);
  if (b->exec)
    (void)hipGraphExecDestroy(b->exec);
  if (b->graph)
    (void)hipGraphDestroy(b->graph);
  (void)hipFree(b->d_subs);
  (void)hipFree(b->d_tasks);
  (void)hipFree(b->d_tasks_node);
  (void)hipFree(b->d_ntasks);
  b->dev->used -= b->bytes;
  delete b;
  return BFAMX_OK;
}

static int batch_any_node(const bfamx_batch_t *b)
{
  for (int n = 0; n < 9; ++n)
    if (b->grp_cnt[n] && node_kernel(b->dev, b->prec, n))
      return 1;
  return 0;
}

static GlueArgs glue_args(const bfamx_batch_t *b)
{
  GlueArgs g;
  g.subs = b->d_subs;
  g.groups = b->glue->d_groups;
  g.ngroups = b->glue->ngroups;
  g.ops = b->glue->d_ops;
  g.buf = b->glue->d_buf;
  return g;
}

/* the order of a batch with the most volume nodes (the one launched on the
 * batch's stream when orders launch side by side) */
static int batch_big_order(const bfamx_batch_t *b)
{
  int big = -1;
  for (int n = 0; n < 9; ++n)
    if (b->grp_cnt[n] &&
        (big < 0 || (size_t)b->grp_cnt[n] * (n + 1) * (n + 1) * (n + 1) >
                        (size_t)b->grp_cnt[big] * (big + 1) * (big + 1) * (big + 1)))
      big = n;
  return big;
}

/* the stage kernels of a batch: one launch (one order, or all orders with
 * the mixed kernel), or one launch per order over its slice of the tasks */
/* the stage kernels of the orders [n] of a batch with their chosen kernel:
 * the line kernel over the line tasks, or the node-parallel kernel over the
 * node tasks */
static hipError_t launch_order(bfamx_batch_t *b, const LaunchArgs &a, int n,
                               hipStream_t st)
{
  LaunchArgs g = a;
  if (node_kernel(b->dev, b->prec, n))
  {
    g.tasks = b->d_tasks_node + b->grpn_off[n];
    g.ntasks = b->grpn_cnt[n];
    /* an order that runs beside the batch's largest keeps whole tasks,
     * in the autotuner's trials as in the batch stage */
    g.whole_tasks = b->N < 0 && n != batch_big_order(b);
    const int slot = kernel_slot(b->dev, b->prec, n);
    return launch_elastic_stage(b->prec, n, g, grid_node(b->dev, b->prec, n, g.ntasks), st,
                                kernel_dyn(b->dev, b->prec, n, slot), slot == 2);
  }
  g.tasks = a.tasks + b->grp_off[n];
  g.ntasks = b->grp_cnt[n];
  return launch_elastic_line(b->prec, n, g, grid_line(b->dev, b->prec, n, g.ntasks), st,
                             kernel_dyn(b->dev, b->prec, n, 0));
}
static hipError_t launch_order(bfamx_batch_t *b, const LaunchArgs &a, int n)
{
  return launch_order(b, a, n, b->dev->stream);
}

/* the stage kernels of a batch: one launch (one order, or all orders with
 * the order-switching kernel), or one launch per order over its slice of
 * the tasks */
static hipError_t launch_batch_kernels(bfamx_batch_t *b, const LaunchArgs &a,
                                       int grid)
{
  if (b->physics != BFAMX_ELASTIC3D)
    return launch_acoustic_stage(b->prec, b->N, a, grid, b->dev->stream);
  if (b->N >= 0)
    return launch_order(b, a, b->N);
  if (!b->grouped)
    return launch_elastic_line(b->prec, b->N, a, grid_line(b->dev, b->prec, b->N, a.ntasks),
                               b->dev->stream);
  /* one launch per order: the largest on the batch's stream, the others on
   * side streams forked after and joined before the next launch, so small
   * orders do not wait for the tail of a large one */
  int ord[9], no = 0;
  const int big = batch_big_order(b);
  for (int n = 0; n < 9; ++n)
    if (b->grp_cnt[n])
      ord[no++] = n;
  bfamx_device_t *d = b->dev;
  hipError_t e = hipSuccess;
  if (no > 1)
    e = hipEventRecord(d->fork_ev, d->stream);
  int k = 0;
  for (int i = 0; i < no && e == hipSuccess; ++i)
  {
    if (ord[i] == big)
      continue;
    e = hipStreamWaitEvent(d->side[k], d->fork_ev, 0);
    if (e == hipSuccess)
      e = launch_order(b, a, ord[i], d->side[k]);
    if (e == hipSuccess)
      e = hipEventRecord(d->join_ev[k], d->side[k]);
    ++k;
  }
  if (e == hipSuccess && big >= 0)
    e = launch_order(b, a, big, d->stream);
  for (int i = 0; i < k && e == hipSuccess; ++i)
    e = hipStreamWaitEvent(d->stream, d->join_ev[i], 0);
  return e;
}

static int batch_stage(bfamx_batch_t *b)
{
  LaunchArgs a;
  memset(&a, 0, sizeof(a));
  a.subs = b->d_subs;
  a.tasks = b->d_tasks;
  a.ntasks = b->ntasks;
  a.mode = MODE_STAGE;
  a.xcd_remap = b->dev->xcd_remap;
  a.nxcd = b->dev->num_xcd;
  const int grid = grid_for(b->dev, b->ntasks);
  hipError_t e = hipSuccess;
  if (b->glue)
    e = launch_mortar(b->prec, glue_args(b), b->dev->stream);
  if (e == hipSuccess)
    e = launch_batch_kernels(b, a, grid);
  if (e == hipSuccess)
    e = launch_clock_advance(b->d_subs, (int)b->subs.size(), b->dev->stream);
  if (e != hipSuccess)
    return bfamx_fail(BFAMX_EHIP, "batch stage launch: %s", hipGetErrorString(e));
  return BFAMX_OK;
}

/* stages of one step: bfamx_batch_create admits one stage count only */
static int batch_nstages(const bfamx_batch_t *b)
{
  return b->subs[0]->h.nstages;
}

/* every member has a state; the members of a glued batch, whose mortars
 * couple their traces at each stage, also share t0, dt and the stage count */
static int batch_ready(const bfamx_batch_t *b)
{
  for (auto *s : b->subs)
    if (!s->fields_set)
      return bfamx_fail(BFAMX_ESTATE, "batch member without state");
  if (b->glue)
  {
    const bfamx_subdomain_t *s0 = b->subs[0];
    for (size_t i = 1; i < b->subs.size(); ++i)
    {
      const bfamx_subdomain_t *s = b->subs[i];
      if (s->t0_set != s0->t0_set || s->dt_set != s0->dt_set ||
          s->stages_run != s0->stages_run)
        return bfamx_fail(BFAMX_ESTATE,
                          "glued batch member %zu is out of step with member 0 "
                          "(t0 %.17g / %.17g, dt %.17g / %.17g, stages %lld / %lld)",
                          i, s->t0_set, s0->t0_set, s->dt_set, s0->dt_set,
                          s->stages_run, s0->stages_run);
    }
  }
  return BFAMX_OK;
}

extern "C" int bfamx_batch_step(bfamx_batch_t *b, int nsteps)
{
  if (!b || nsteps < 0)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_batch_step: bad arguments");
  int rc = batch_ready(b);
  if (rc)
    return rc;
  const int ns = batch_nstages(b);
  for (int i = 0; i < nsteps * ns; ++i)
  {
    rc = batch_stage(b);
    if (rc)
      return rc;
    for (auto *s : b->subs)
      ++s->stages_run;
  }
  return BFAMX_OK;
}

extern "C" int bfamx_batch_capture(bfamx_batch_t *b, int nsteps)
{
  if (!b || nsteps < 1)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_batch_capture: bad arguments");
  if (b->exec)
  {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = NULL;
  }
  if (b->graph)
  {
    (void)hipGraphDestroy(b->graph);
    b->graph = NULL;
  }
  hipStream_t st = b->dev->stream;
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const int ns = batch_nstages(b);
  int rc = BFAMX_OK;
  for (int i = 0; i < nsteps * ns && rc == BFAMX_OK; ++i)
    rc = batch_stage(b);
  hipGraph_t g = NULL;
  hipError_t e = hipStreamEndCapture(st, &g);
  if (rc)
  {
    if (g)
      (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess)
    return bfamx_fail(BFAMX_EHIP, "graph capture: %s", hipGetErrorString(e));
  b->graph = g;
  HIPCHK(hipGraphInstantiate(&b->exec, g, NULL, NULL, 0));
  b->graph_steps = nsteps;
  return BFAMX_OK;
}

extern "C" int bfamx_batch_launch_graph(bfamx_batch_t *b)
{
  if (!b || !b->exec)
    return bfamx_fail(BFAMX_ESTATE, "bfamx_batch_launch_graph: not captured");
  int rc = batch_ready(b);
  if (rc)
    return rc;
  HIPCHK(hipGraphLaunch(b->exec, b->dev->stream));
  for (auto *s : b->subs)
    s->stages_run += (long long)b->graph_steps * batch_nstages(b);
  return BFAMX_OK;
}

/* ---------------------------------------------------------------------- */
/* glue: mortars between subdomains of one batch                           */
/* ---------------------------------------------------------------------- */

static int glue_release(bfamx_glue_t *g)
{
  /* restore the face tables of the glued faces: one download, patch and
   * upload per subdomain */
  for (size_t i = 0; i < g->subs.size(); ++i)
  {
    bfamx_subdomain_t *s = g->subs[i];
    bool any = false;
    for (const auto &f : g->saved)
      any = any || f.sub == (int)i;
    if (!any)
      continue;
    const size_t nfc = (size_t)s->K * 6;
    std::vector<int32_t> ne(nfc);
    std::vector<uint8_t> nc(nfc);
    if (hipMemcpy(ne.data(), s->h.nbr_elem, 4 * nfc, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(nc.data(), s->h.nbr_code, nfc, hipMemcpyDeviceToHost) != hipSuccess)
      continue;
    for (const auto &f : g->saved)
      if (f.sub == (int)i)
      {
        ne[f.idx] = f.elem;
        nc[f.idx] = f.code;
      }
    (void)hipMemcpy((void *)s->h.nbr_elem, ne.data(), 4 * nfc, hipMemcpyHostToDevice);
    (void)hipMemcpy((void *)s->h.nbr_code, nc.data(), nfc, hipMemcpyHostToDevice);
  }
  for (auto *s : g->subs)
    if (s->glue == g)
    {
      s->glue = NULL;
      s->h.glue_buf = NULL;
      (void)push_sub(s);
    }
  dev_free_all(g->dev, g->allocs, g->bytes);
  delete g;
  return BFAMX_OK;
}

extern "C" int bfamx_glue_free(bfamx_glue_t *g)
{
  if (!g)
    return BFAMX_OK;
  (void)hipStreamSynchronize(g->dev->stream);
  return glue_release(g);
}

template <class T>
static std::vector<T> glue_ops_table()
{
  std::vector<T> t(GLUE_OPS_SIZE, T(0));
  double prj[5 * 81], mp[5 * 81], wmp[5 * 81];
  for (int Na = 1; Na <= 8; ++Na)
    for (int Nb = 1; Nb <= 8; ++Nb)
    {
      int ident = 0;
      if (bfamx_mortar_ops(Na, Nb, prj, mp, wmp, &ident))
        continue;
      const int blk = (Na + 1) * (Nb + 1);
      for (int k = 0; k < 5; ++k)
        for (int i = 0; i < blk; ++i)
        {
          t[glue_ops_index(0, Na, Nb, k) + i] = (T)prj[k * blk + i];
          t[glue_ops_index(1, Na, Nb, k) + i] = (T)wmp[k * blk + i];
        }
    }
  return t;
}

/* the mortar kernel's operators: glue_ops_table's blocks between orders
 * 1..7, padded to 8 x 8 */
template <class T>
static std::vector<T> glue_ops8_table()
{
  const std::vector<T> t = glue_ops_table<T>();
  std::vector<T> p(GLUE_OPS8_SIZE, T(0));
  for (int Na = 1; Na <= 7; ++Na)
    for (int Nb = 1; Nb <= 7; ++Nb)
      for (int kind = 0; kind < 2; ++kind)
        for (int k = 0; k < 5; ++k)
        {
          /* (Nb + 1) x (Na + 1), column-major */
          const T *src = t.data() + glue_ops_index(kind, Na, Nb, k);
          T *dst = p.data() + glue_ops8_index(kind, Na, Nb, k);
          for (int c = 0; c <= Na; ++c)
            for (int r = 0; r <= Nb; ++r)
              dst[r + 8 * c] = src[r + (Nb + 1) * c];
        }
  return p;
}

extern "C" int bfamx_glue_upload(bfamx_device_t *dev, bfamx_subdomain_t *const *subs,
                                 int nsubs, const bfamx_mortar_t *m, int32_t nm,
                                 bfamx_glue_t **out)
{
  if (!dev || !subs || nsubs < 1 || nm < 0 || (nm && !m) || !out)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_glue_upload: bad arguments");
  *out = NULL;
  const int prec = subs[0]->prec;
  for (int i = 0; i < nsubs; ++i)
  {
    const bfamx_subdomain_t *s = subs[i];
    if (!s || s->dev != dev || s->physics != BFAMX_ELASTIC3D || s->prec != prec ||
        prec == BFAMX_FP64_PLAIN || s->N < 3 || s->N > 7)
      return bfamx_fail(BFAMX_EINVAL,
                        "glue: elastic fused subdomains of one precision, orders 3..7");
    if (s->glue)
      return bfamx_fail(BFAMX_ESTATE, "glue: subdomain %d is already glued", i);
    /* the mortars couple the members' traces at every stage: one table */
    const bfamx_lsrk_t &a = s->lsrk, &b0 = subs[0]->lsrk;
    const size_t ns = (size_t)b0.n_stages;
    if (a.n_stages != b0.n_stages || memcmp(a.A, b0.A, ns * sizeof(double)) ||
        memcmp(a.B, b0.B, ns * sizeof(double)) || memcmp(a.C, b0.C, (ns + 1) * sizeof(double)))
      return bfamx_fail(BFAMX_EINVAL,
                        "glue: subdomain %d uses another LSRK table than subdomain 0", i);
  }
  /* group the mortars by b face: one conforming mortar, or four quadrants */
  std::vector<int32_t> ord(nm);
  for (int32_t i = 0; i < nm; ++i)
    ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) {
    const bfamx_mortar_t &a = m[x], &b = m[y];
    if (a.b_sub != b.b_sub)
      return a.b_sub < b.b_sub;
    if (a.b_elem != b.b_elem)
      return a.b_elem < b.b_elem;
    if (a.b_face != b.b_face)
      return a.b_face < b.b_face;
    return a.quad < b.quad;
  });
  for (int32_t i = 0; i < nm; ++i)
  {
    const bfamx_mortar_t &r = m[i];
    if (r.s_sub < 0 || r.s_sub >= nsubs || r.b_sub < 0 || r.b_sub >= nsubs ||
        r.s_elem < 0 || r.s_elem >= subs[r.s_sub]->K || r.b_elem < 0 ||
        r.b_elem >= subs[r.b_sub]->K || r.s_face < 0 || r.s_face > 5 ||
        r.b_face != (r.s_face ^ 1) || r.quad < -1 || r.quad > 3)
      return bfamx_fail(BFAMX_EINVAL, "glue: mortar %d is malformed", i);
    const int nmax = std::max(subs[r.s_sub]->N, subs[r.b_sub]->N);
    if (r.order != 0 && (r.order < nmax || r.order > 7))
      return bfamx_fail(BFAMX_EINVAL,
                        "glue: mortar %d has order %d: 0, or %d..7 (at least both sides' "
                        "orders)", i, (int)r.order, nmax);
  }
  std::vector<MortarGroup> groups;
  /* face -> offset in the glue buffer; every glued face appears once */
  std::vector<std::vector<int32_t>> off(nsubs);
  for (int i = 0; i < nsubs; ++i)
    off[i].assign((size_t)subs[i]->K * 6, -1);
  size_t total = 0;
  auto slot = [&](int sub, int32_t e, int f, int32_t *o) -> int {
    int32_t &x = off[sub][(size_t)e * 6 + f];
    if (x >= 0)
      return bfamx_fail(BFAMX_EINVAL, "glue: face %d of element %d (subdomain %d) is "
                        "named by more than one interface", f, e, sub);
    const int n1 = subs[sub]->N + 1;
    if (total + 6 * (size_t)n1 * n1 > (size_t)INT32_MAX)
      return bfamx_fail(BFAMX_EINVAL, "glue: too many glued faces");
    x = (int32_t)total;
    total += 6 * (size_t)n1 * n1;
    *o = x;
    return BFAMX_OK;
  };
  int rc = BFAMX_OK;
  for (int32_t i = 0; i < nm && !rc;)
  {
    const bfamx_mortar_t &r = m[ord[i]];
    MortarGroup g;
    memset(&g, 0, sizeof(g));
    g.b_sub = r.b_sub;
    g.b_elem = r.b_elem;
    g.b_face = r.b_face;
    int n = 1;
    if (r.quad >= 0)
    {
      n = 4;
      for (int k = 0; k < 4; ++k)
      {
        if (i + k >= nm)
          return bfamx_fail(BFAMX_EINVAL, "glue: hanging face with fewer than 4 mortars");
        const bfamx_mortar_t &q = m[ord[i + k]];
        if (q.b_sub != r.b_sub || q.b_elem != r.b_elem || q.b_face != r.b_face ||
            q.quad != k)
          return bfamx_fail(BFAMX_EINVAL, "glue: hanging face without quadrants 0..3");
      }
    }
    g.nm = (int8_t)n;
    g.quad0 = n == 1 ? -1 : 0;
    if ((rc = slot(r.b_sub, r.b_elem, r.b_face, &g.b_off)))
      break;
    for (int k = 0; k < n && !rc; ++k)
    {
      const bfamx_mortar_t &q = m[ord[i + k]];
      g.s_sub[k] = q.s_sub;
      g.s_elem[k] = q.s_elem;
      g.s_face[k] = q.s_face;
      g.s_ng[k] = q.order;
      rc = slot(q.s_sub, q.s_elem, q.s_face, &g.s_off[k]);
    }
    groups.push_back(g);
    i += n;
  }
  if (rc)
    return rc;
  const int fp32 = prec != BFAMX_FP64;
  bfamx_glue_t *gl = new bfamx_glue_t();
  gl->dev = dev;
  gl->subs.assign(subs, subs + nsubs);
  gl->prec = prec;
  gl->ngroups = (int)groups.size();
  gl->bytes = 0;
  void *p;
  rc = dev_alloc(dev, gl->allocs, &gl->bytes, sizeof(MortarGroup) * (groups.size() + 1), &p);
  if (!rc)
  {
    gl->d_groups = (MortarGroup *)p;
    rc = dev_alloc(dev, gl->allocs, &gl->bytes, (fp32 ? 4 : 8) * (size_t)GLUE_OPS8_SIZE, &p);
  }
  if (!rc)
  {
    gl->d_ops = p;
    rc = dev_alloc(dev, gl->allocs, &gl->bytes, (fp32 ? 4 : 8) * (total + 1), &p);
  }
  if (rc)
  {
    dev_free_all(dev, gl->allocs, gl->bytes);
    delete gl;
    return rc;
  }
  gl->d_buf = p;
  hipError_t he = hipMemcpy(gl->d_groups, groups.data(), sizeof(MortarGroup) * groups.size(),
                            hipMemcpyHostToDevice);
  if (he == hipSuccess)
  {
    if (fp32)
    {
      auto t = glue_ops8_table<float>();
      he = hipMemcpy(gl->d_ops, t.data(), 4 * t.size(), hipMemcpyHostToDevice);
    }
    else
    {
      auto t = glue_ops8_table<double>();
      he = hipMemcpy(gl->d_ops, t.data(), 8 * t.size(), hipMemcpyHostToDevice);
    }
  }
  if (he == hipSuccess)
    he = hipMemset(gl->d_buf, 0, (fp32 ? 4 : 8) * (total + 1));
  /* flag the glued faces in the subdomains' face tables */
  for (int i = 0; i < nsubs && he == hipSuccess; ++i)
  {
    bfamx_subdomain_t *s = subs[i];
    const size_t nfc = (size_t)s->K * 6;
    std::vector<int32_t> ne(nfc);
    std::vector<uint8_t> nc(nfc);
    he = hipMemcpy(ne.data(), s->h.nbr_elem, 4 * nfc, hipMemcpyDeviceToHost);
    if (he == hipSuccess)
      he = hipMemcpy(nc.data(), s->h.nbr_code, nfc, hipMemcpyDeviceToHost);
    if (he != hipSuccess)
      break;
    for (size_t k = 0; k < nfc; ++k)
      if (off[i][k] >= 0)
      {
        gl->saved.push_back({i, (int32_t)k, ne[k], nc[k]});
        ne[k] = off[i][k];
        nc[k] = (uint8_t)NBR_GLUED;
      }
    he = hipMemcpy((void *)s->h.nbr_elem, ne.data(), 4 * nfc, hipMemcpyHostToDevice);
    if (he == hipSuccess)
      he = hipMemcpy((void *)s->h.nbr_code, nc.data(), nfc, hipMemcpyHostToDevice);
    s->glue = gl;
    s->h.glue_buf = gl->d_buf;
    if (he == hipSuccess && push_sub(s))
      he = hipErrorUnknown;
  }
  if (he != hipSuccess)
  {
    glue_release(gl);
    return bfamx_fail(BFAMX_EHIP, "glue upload failed: %s", hipGetErrorString(he));
  }
  *out = gl;
  return BFAMX_OK;
}

extern "C" int bfamx_batch_set_glue(bfamx_batch_t *b, bfamx_glue_t *g)
{
  if (!b)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_batch_set_glue: NULL batch");
  if (g)
  {
    if (g->subs.size() != b->subs.size())
      return bfamx_fail(BFAMX_EINVAL, "glue and batch must hold the same subdomains");
    for (size_t i = 0; i < b->subs.size(); ++i)
      if (g->subs[i] != b->subs[i])
        return bfamx_fail(BFAMX_EINVAL, "glue and batch must list the subdomains in "
                          "the same order");
  }
  if (b->exec)
    return bfamx_fail(BFAMX_ESTATE, "bfamx_batch_set_glue: batch already captured");
  b->glue = g;
  /* refresh the batch's descriptor copies (glue buffer, face flags) */
  std::vector<SubDev> hs(b->subs.size());
  for (size_t i = 0; i < b->subs.size(); ++i)
    hs[i] = b->subs[i]->h;
  HIPCHK(hipMemcpy(b->d_subs, hs.data(), sizeof(SubDev) * hs.size(), hipMemcpyHostToDevice));
  return BFAMX_OK;
}

extern "C" int bfamx_batch_rhs(bfamx_batch_t *b, double t, double *const *const *rhs)
{
  if (!b || !rhs)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_batch_rhs: bad arguments");
  for (auto *s : b->subs)
    if (!s->fields_set)
      return bfamx_fail(BFAMX_ESTATE, "batch member without state");
  if (b->glue)
  {
    hipError_t e = launch_mortar(b->prec, glue_args(b), b->dev->stream);
    if (e != hipSuccess)
      return bfamx_fail(BFAMX_EHIP, "mortar launch: %s", hipGetErrorString(e));
  }
  for (size_t i = 0; i < b->subs.size(); ++i)
  {
    int rc = rhs_one(b->subs[i], t, rhs[i]);
    if (rc)
      return rc;
  }
  return BFAMX_OK;
}

/* ---------------------------------------------------------------------- */
/* autotuning of the mixed-order launch                                    */
/* ---------------------------------------------------------------------- */

extern "C" int bfamx_batch_set_strategy(bfamx_batch_t *b, int per_order)
{
  if (!b || per_order < 0 || per_order > 1)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_batch_set_strategy: bad arguments");
  if (b->exec)
    return bfamx_fail(BFAMX_ESTATE, "bfamx_batch_set_strategy: batch already captured");
  b->grouped = b->N < 0 ? per_order : 0;
  return BFAMX_OK;
}

/*
 * Autotuning on the batch's own data.  For every order present it times
 * `trials` stage launches of the line kernel and of the node-parallel
 * kernel and keeps the faster; with mixed orders it then times one
 * order-switching launch against one launch per order.  The trial launches
 * do not advance the clocks: they read the current state, write the free
 * ping-pong buffer and dq, and dq is saved beforehand and restored, so the
 * batch's state and results are unchanged.  Choices are cached on the
 * device (per precision and order, and per precision and set of orders).
 */
extern "C" int bfamx_batch_autotune(bfamx_batch_t *b, int trials, int force,
                                    int *choice)
{
  if (!b || trials < 1)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_batch_autotune: bad arguments");
  for (auto *s : b->subs)
    if (!s->fields_set)
      return bfamx_fail(BFAMX_ESTATE, "batch member without state");
  if (b->physics != BFAMX_ELASTIC3D)
  {
    if (choice)
      *choice = 0;
    return BFAMX_OK;
  }
  if (b->exec)
    return bfamx_fail(BFAMX_ESTATE, "bfamx_batch_autotune: batch already captured");
  bfamx_device_t *d = b->dev;
  signed char *kc = d->kchoice[b->prec];
  signed char &mc = d->mixed_choice[b->prec][b->order_mask & 255];
  int need = 0;
  for (int n = 0; n < 9; ++n)
    if (b->grp_cnt[n] && (kc[n] < 0 || force))
      need = 1;
  if (b->N < 0 && (mc < 0 || force))
    need = 1;
  if (!need)
  {
    b->grouped = b->N < 0 ? (mc > 0 || batch_any_node(b)) : 0;
    if (choice)
      *choice = b->grouped;
    return BFAMX_OK;
  }
  /* the trial launches repeat the current stage without advancing the
   * clock: they rewrite the next state's buffer and dq.  Save dq, or, where
   * dq is rebuilt from the states, the previous state that the next
   * buffer holds until the stage ends */
  std::vector<void *> saved(b->subs.size(), nullptr), reg(b->subs.size(), nullptr);
  std::vector<void *> allocs;
  size_t bytes = 0;
  int rc = BFAMX_OK;
  for (size_t i = 0; i < b->subs.size() && !rc; ++i)
  {
    bfamx_subdomain_t *s = b->subs[i];
    const size_t nb = (size_t)s->K * s->nf * s->Np * (s->rate_fp32 ? 4 : 8);
    reg[i] = s->h.dq;
    if (s->h.dq_rec)
    {
      int parity = 0;
      rc = current_parity(s, &parity);
      reg[i] = s->h.q[parity ^ 1];
    }
    if (!rc)
      rc = dev_alloc(d, allocs, &bytes, nb, &saved[i]);
    if (!rc && hipMemcpyAsync(saved[i], reg[i], nb, hipMemcpyDeviceToDevice, d->stream) !=
                   hipSuccess)
      rc = bfamx_fail(BFAMX_EHIP, "autotune: dq save failed");
  }
  LaunchArgs a;
  memset(&a, 0, sizeof(a));
  a.subs = b->d_subs;
  a.tasks = b->d_tasks;
  a.ntasks = b->ntasks;
  a.mode = MODE_STAGE;
  a.xcd_remap = d->xcd_remap;
  a.nxcd = d->num_xcd;
  /* each trial launch starts from the saved buffers: a launch rewrites
   * what the next one would read (dq, or the previous state that the
   * rebuilt dq comes from), and without the restore the timed data would
   * grow by |A_s| (or |B_s A_s / B_{s-1}|) per launch */
  auto restore = [&]() -> hipError_t {
    for (size_t i = 0; i < b->subs.size(); ++i)
      if (saved[i])
      {
        bfamx_subdomain_t *s = b->subs[i];
        const size_t nb = (size_t)s->K * s->nf * s->Np * (s->rate_fp32 ? 4 : 8);
        hipError_t e = hipMemcpyAsync(reg[i], saved[i], nb, hipMemcpyDeviceToDevice, d->stream);
        if (e != hipSuccess)
          return e;
      }
    return hipSuccess;
  };
  /* mean time of `trials` launches of fn() after one warm-up, each timed
   * alone between events after the restore */
  auto time_it = [&](auto fn, float *ms) -> int {
    hipError_t e = restore();
    if (e == hipSuccess)
      e = fn();
    float total = 0.f;
    for (int t = 0; t < trials && e == hipSuccess; ++t)
    {
      float one = 0.f;
      e = restore();
      if (e == hipSuccess)
        e = hipEventRecord(d->ev0, d->stream);
      if (e == hipSuccess)
        e = fn();
      if (e == hipSuccess)
        e = hipEventRecord(d->ev1, d->stream);
      if (e == hipSuccess)
        e = hipEventSynchronize(d->ev1);
      if (e == hipSuccess)
        e = hipEventElapsedTime(&one, d->ev0, d->ev1);
      total += one;
    }
    *ms = total;
    return e == hipSuccess ? BFAMX_OK
                           : bfamx_fail(BFAMX_EHIP, "autotune launch: %s", hipGetErrorString(e));
  };
  /* (a) kernel, block size and resident-block cap per order: the line
   * kernel, the node-parallel kernel with one block per task and on a
   * resident grid, each with both its block sizes at N = 3..5 (kinds 3, 4),
   * each uncapped and capped at 3/4 and 1/2 of its resident blocks per CU
   * (split-BF16x3 has the node-parallel kernel only); caps the occupancy
   * calculator cannot place are skipped */
  for (int n = 0; n < 9 && !rc; ++n)
  {
    if (!b->grp_cnt[n] || (kc[n] >= 0 && !force))
      continue;
    float best = 0.f;
    int bk = -1, bl = 0;
    const int kmax = n >= 3 && n <= 5 ? 5 : 3;
    for (int k = b->prec == PREC_FP32_BF16X3 ? 1 : 0; k < kmax && !rc; ++k)
      for (int lv = 0; lv < 3 && !rc; ++lv)
      {
        const int slot = kind_slot(k);
        if (lv > 0 && occ_dyn(d, b->prec, n, slot, lv) == 0)
          continue;
        kc[n] = (signed char)k;
        d->occ[b->prec][n][slot] = (signed char)lv;
        float ms = 0.f;
        rc = time_it([&]() { return launch_order(b, a, n); }, &ms);
        if (!rc && (bk < 0 || ms < best))
          best = ms, bk = k, bl = lv;
      }
    for (int sl = 0; sl < 3; ++sl)
      d->occ[b->prec][n][sl] = 0;
    if (rc)
    {
      kc[n] = -1;
      break;
    }
    kc[n] = (signed char)bk;
    d->occ[b->prec][n][kind_slot(bk)] = (signed char)bl;
  }
  /* (b) one order-switching launch or one launch per order.  The
   * order-switching launch is the line kernel: it is a candidate only when
   * every order of the batch runs the line kernel, so that each element is
   * computed by the same code as in a single-subdomain launch */
  const int keep = b->grouped;
  if (!rc && b->N < 0 && (mc < 0 || force) && batch_any_node(b))
    mc = 1;
  if (!rc && b->N < 0 && (mc < 0 || force))
  {
    float ms[2] = {0, 0};
    for (int mode = 0; mode < 2 && !rc; ++mode)
    {
      b->grouped = mode;
      rc = time_it([&]() { return launch_batch_kernels(b, a, grid_for(d, b->ntasks)); },
                   &ms[mode]);
    }
    if (!rc)
      mc = (signed char)(ms[1] < ms[0] ? 1 : 0);
  }
  if (restore() != hipSuccess && !rc)
    rc = bfamx_fail(BFAMX_EHIP, "autotune: dq restore failed");
  (void)hipStreamSynchronize(d->stream);
  dev_free_all(d, allocs, bytes);
  if (rc)
  {
    b->grouped = keep;
    return rc;
  }
  b->grouped = b->N < 0 ? (mc > 0 || batch_any_node(b)) : 0;
  if (choice)
    *choice = b->grouped;
  return BFAMX_OK;
}

extern "C" int bfamx_set_kernel(bfamx_device_t *d, int prec, int N, int kind)
{
  if (!d || prec < 0 || prec >= NPREC || N < 0 || N > 8 || kind < -1 || kind > 4)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_set_kernel: bad arguments");
  if (prec == PREC_FP32_BF16X3 && kind == 0)
    return bfamx_fail(BFAMX_EUNSUPPORTED,
                      "bfamx_set_kernel: split-BF16x3 runs the node-parallel kernel only");
  d->kchoice[prec][N] = (signed char)kind;
  return BFAMX_OK;
}

extern "C" int bfamx_set_occupancy(bfamx_device_t *d, int prec, int N, int kind, int level)
{
  if (!d || prec < 0 || prec >= NPREC || N < 3 || N > 7 || kind < 0 || kind > 2 ||
      level < 0 || level > 2)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_set_occupancy: bad arguments");
  d->occ[prec][N][kind] = (signed char)level;
  return BFAMX_OK;
}

extern "C" int bfamx_get_occupancy(const bfamx_device_t *d, int prec, int N, int kind,
                                   int *level)
{
  if (!d || prec < 0 || prec >= NPREC || N < 3 || N > 7 || kind < 0 || kind > 2 || !level)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_get_occupancy: bad arguments");
  *level = d->occ[prec][N][kind];
  return BFAMX_OK;
}

extern "C" int bfamx_get_kernel(const bfamx_device_t *d, int prec, int N, int *kind)
{
  if (!d || prec < 0 || prec >= NPREC || N < 0 || N > 8 || !kind)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_get_kernel: bad arguments");
  *kind = kernel_kind(d, prec, N);
  return BFAMX_OK;
}

static int grid_line(bfamx_device_t *d, int prec, int N, int ntasks)
{
  const int nx = d->xcd_remap == 1 && d->num_xcd > 1 ? d->num_xcd : 1;
  int g = nx * ((ntasks + nx - 1) / nx);
  if (!d->persistent)
    return g;
  const int ni = N < 0 ? 8 : N;
  int &bpc = d->line_bpc[prec][ni];
  if (bpc == 0)
  {
    int b = 0;
    if (line_blocks_per_cu(prec, N, &b) != hipSuccess || b < 1)
      b = -1;
    bpc = b;
  }
  if (bpc < 0)
    return g;
  const int resident = nx * ((d->num_cu * bpc + nx - 1) / nx);
  return resident < g ? resident : g;
}

extern "C" int bfamx_set_persistent(bfamx_device_t *d, int on)
{
  if (!d || on < 0 || on > 1)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_set_persistent: bad arguments");
  d->persistent = on;
  return BFAMX_OK;
}

/* ---------------------------------------------------------------------- */
/* physical boundaries                                                     */
/* ---------------------------------------------------------------------- */

extern "C" int bfamx_set_boundary(bfamx_subdomain_t *s, const int8_t *kind)
{
  if (!s || !kind)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_set_boundary: NULL argument");
  if (s->glue)
    return bfamx_fail(BFAMX_ESTATE, "bfamx_set_boundary: set boundaries before gluing");
  const int nfaces = 2 * s->dim;
  const size_t nfc = (size_t)s->K * nfaces;
  std::vector<int32_t> ne(nfc);
  std::vector<uint8_t> nc(nfc);
  HIPCHK(hipStreamSynchronize(s->dev->stream));
  HIPCHK(hipMemcpy(ne.data(), s->h.nbr_elem, 4 * nfc, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(nc.data(), s->h.nbr_code, nfc, hipMemcpyDeviceToHost));
  for (size_t k = 0; k < nfc; ++k)
  {
    const int e = (int)(k / nfaces), f = (int)(k % nfaces);
    const int b = kind[k];
    if (b < 0 || b > 3)
      return bfamx_fail(BFAMX_EINVAL, "bfamx_set_boundary: kind %d at face %d of %d", b, f, e);
    const int self = ne[k] == e && (nc[k] & 7) == f && !(nc[k] & NBR_GLUED);
    if (b == 0)
    {
      if (nc[k] & NBR_BOUNDARY)
        nc[k] = (uint8_t)f; /* back to the plain self face */
      continue;
    }
    if (!self)
      return bfamx_fail(BFAMX_EINVAL,
                        "bfamx_set_boundary: face %d of element %d has a neighbour", f, e);
    nc[k] = (uint8_t)(NBR_BOUNDARY | (b << 3) | f);
  }
  HIPCHK(hipMemcpy((void *)s->h.nbr_code, nc.data(), nfc, hipMemcpyHostToDevice));
  return BFAMX_OK;
}

/* ---------------------------------------------------------------------- */
/* adaptation: field transfer between meshes                               */
/* ---------------------------------------------------------------------- */

extern "C" int bfamx_transfer_fields(const bfamx_transfer_maps_t *m,
                                     bfamx_subdomain_t *const *src, int nsrc,
                                     bfamx_subdomain_t *const *dst, int ndst,
                                     const int32_t *dst_subd, const int32_t *dst_elem,
                                     double dt)
{
  if (!m || !src || nsrc < 1 || !dst || ndst < 1 || !dst_subd || !dst_elem ||
      m->num_dst < 0 || (m->num_dst && (!m->dst_to_adapt_flags || !m->dst_to_dst_chld_id ||
                                        !m->dst_to_src_subd_id || !m->dst_to_src_elem_id)) ||
      m->num_coarse_dst < 0 ||
      (m->num_coarse_dst && (!m->coarse_dst_to_src_chld_id || !m->coarse_dst_to_src_subd_id ||
                             !m->coarse_dst_to_src_elem_id)))
    return bfamx_fail(BFAMX_EINVAL, "bfamx_transfer_fields: bad arguments");
  bfamx_device_t *dev = src[0]->dev;
  const int sfp32 = src[0]->state_fp32, dfp32 = dst[0]->state_fp32;
  for (int i = 0; i < nsrc + ndst; ++i)
  {
    const bfamx_subdomain_t *s = i < nsrc ? src[i] : dst[i - nsrc];
    if (!s || s->dev != dev || s->physics != BFAMX_ELASTIC3D ||
        s->state_fp32 != (i < nsrc ? sfp32 : dfp32))
      return bfamx_fail(BFAMX_EINVAL, "transfer: elastic subdomains of one device, one "
                                      "state type among the sources and one among the "
                                      "destinations");
    if (s->glue)
      return bfamx_fail(BFAMX_ESTATE, "transfer: free the glue of subdomain %d first", i);
  }
  /* the sources stand between steps, at one time */
  double t = 0;
  for (int i = 0; i < nsrc; ++i)
  {
    if (!src[i]->fields_set)
      return bfamx_fail(BFAMX_ESTATE, "transfer: source %d has no state", i);
    double ti;
    int64_t st;
    int rc = bfamx_get_time(src[i], &ti, &st);
    if (rc)
      return rc;
    if (st % src[i]->h.nstages)
      return bfamx_fail(BFAMX_ESTATE, "transfer: source %d is inside a step (stage %lld)", i,
                        (long long)st);
    if (i == 0)
      t = ti;
    else if (ti != t)
      return bfamx_fail(BFAMX_ESTATE, "transfer: sources at different times (%.17g, %.17g)",
                        t, ti);
  }
  /* tasks, with every index checked against the subdomains */
  const int32_t nd = m->num_dst;
  std::vector<XferTask> tasks(nd);
  std::vector<std::vector<char>> seen(ndst);
  int32_t covered = 0;
  for (int i = 0; i < ndst; ++i)
    seen[i].assign((size_t)dst[i]->K, 0);
  auto src_ok = [&](int32_t s, int32_t e) { return s >= 0 && s < nsrc && e >= 0 && e < src[s]->K; };
  for (int32_t k = 0; k < nd; ++k)
  {
    XferTask &t = tasks[k];
    memset(&t, 0, sizeof(t));
    const int32_t ds = dst_subd[k], de = dst_elem[k];
    if (ds < 0 || ds >= ndst || de < 0 || de >= dst[ds]->K || seen[ds][de])
      return bfamx_fail(BFAMX_EINVAL, "transfer: destination %d names (%d, %d): out of range "
                                      "or named twice", k, ds, de);
    seen[ds][de] = 1;
    ++covered;
    t.dst_sub = ds;
    t.dst_elem = de;
    const int flag = m->dst_to_adapt_flags[k], cid = m->dst_to_dst_chld_id[k];
    if (flag == BFAMX_FLAG_COARSEN)
    {
      const int32_t c = -1 - m->dst_to_src_subd_id[k];
      if (c < 0 || c >= m->num_coarse_dst)
        return bfamx_fail(BFAMX_EINVAL, "transfer: destination %d: bad coarse index", k);
      t.nsrc = 8;
      for (int j = 0; j < 8; ++j)
      {
        const int32_t s = m->coarse_dst_to_src_subd_id[8 * c + j];
        const int32_t e = m->coarse_dst_to_src_elem_id[8 * c + j];
        const int ch = m->coarse_dst_to_src_chld_id[8 * c + j];
        if (!src_ok(s, e) || ch < 0 || ch > 7)
          return bfamx_fail(BFAMX_EINVAL, "transfer: destination %d: bad child %d", k, j);
        t.src_sub[j] = s;
        t.src_elem[j] = e;
        for (int d = 0; d < 3; ++d)
          t.op[j] |= (1 + ((ch >> d) & 1)) << (4 * d);
      }
    }
    else if (flag == BFAMX_FLAG_SAME || flag == BFAMX_FLAG_REFINE)
    {
      const int32_t s = m->dst_to_src_subd_id[k], e = m->dst_to_src_elem_id[k];
      if (!src_ok(s, e) || cid < 0 || cid > 7)
        return bfamx_fail(BFAMX_EINVAL, "transfer: destination %d: bad source", k);
      t.nsrc = 1;
      t.src_sub[0] = s;
      t.src_elem[0] = e;
      for (int d = 0; d < 3; ++d)
        t.op[0] |= (flag == BFAMX_FLAG_SAME ? 0 : 3 + ((cid >> d) & 1)) << (4 * d);
    }
    else
      return bfamx_fail(BFAMX_EINVAL, "transfer: destination %d has flag %d", k, flag);
  }
  int64_t ktot = 0;
  for (int i = 0; i < ndst; ++i)
    ktot += dst[i]->K;
  if (covered != ktot)
    return bfamx_fail(BFAMX_EINVAL, "transfer: the maps name %d of the %lld destination "
                                    "elements", covered, (long long)ktot);
  /* sides: the sources' current states; the destinations' first buffer */
  std::vector<XferSide> hs(nsrc), hd(ndst);
  for (int i = 0; i < nsrc; ++i)
  {
    int parity = 0, rc = current_parity(src[i], &parity);
    if (rc)
      return rc;
    hs[i].q = src[i]->h.q[parity];
    hs[i].N = src[i]->N;
    hs[i].K = src[i]->K;
  }
  for (int i = 0; i < ndst; ++i)
  {
    hd[i].q = dst[i]->h.q[0];
    hd[i].N = dst[i]->N;
    hd[i].K = dst[i]->K;
  }
  const std::vector<double> ops = glue_ops_table<double>();
  std::vector<void *> allocs;
  size_t bytes = 0;
  void *p_src = NULL, *p_dst = NULL, *p_task = NULL, *p_ops = NULL;
  int rc = dev_alloc(dev, allocs, &bytes, sizeof(XferSide) * nsrc, &p_src);
  if (!rc)
    rc = dev_alloc(dev, allocs, &bytes, sizeof(XferSide) * ndst, &p_dst);
  if (!rc)
    rc = dev_alloc(dev, allocs, &bytes, sizeof(XferTask) * (nd ? nd : 1), &p_task);
  if (!rc)
    rc = dev_alloc(dev, allocs, &bytes, sizeof(double) * ops.size(), &p_ops);
  hipError_t e = hipSuccess;
  if (!rc)
  {
    /* copies on the kernel's stream: ordered before it; the host arrays live
     * until the synchronisation below */
    hipStream_t st = dev->stream;
    e = hipMemcpyAsync(p_src, hs.data(), sizeof(XferSide) * nsrc, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(p_dst, hd.data(), sizeof(XferSide) * ndst, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nd)
      e = hipMemcpyAsync(p_task, tasks.data(), sizeof(XferTask) * nd, hipMemcpyHostToDevice,
                         st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(p_ops, ops.data(), sizeof(double) * ops.size(),
                         hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
      e = launch_transfer(sfp32, dfp32, (const XferSide *)p_src, (const XferSide *)p_dst,
                          (const XferTask *)p_task, nd, (const double *)p_ops, dev->stream);
    if (e == hipSuccess)
      e = hipStreamSynchronize(dev->stream);
    if (e != hipSuccess)
    {
      (void)hipStreamSynchronize(dev->stream); /* nothing pending on the buffers */
      rc = bfamx_fail(BFAMX_EHIP, "transfer: %s", hipGetErrorString(e));
    }
  }
  dev_free_all(dev, allocs, bytes);
  if (rc)
    return rc;
  /* the destinations start a step at t with step dt */
  for (int i = 0; i < ndst; ++i)
  {
    bfamx_subdomain_t *s = dst[i];
    const size_t ndof = (size_t)s->K * s->nf * s->Np;
    if (s->h.dq)
      HIPCHK(hipMemset(s->h.dq, 0, (s->rate_fp32 ? 4 : 8) * ndof));
    Clock c;
    memset(&c, 0, sizeof(c));
    c.t = t;
    c.dt = dt;
    HIPCHK(hipMemcpy(s->d_clock, &c, sizeof(c), hipMemcpyHostToDevice));
    s->fields_set = 1;
    s->t0_set = t;
    s->dt_set = dt;
    s->stages_run = 0;
    rc = refresh_clock(s);
    if (rc)
      return rc;
  }
  return BFAMX_OK;
}
