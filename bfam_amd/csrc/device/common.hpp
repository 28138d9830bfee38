/*
 * Device-side data layout shared by the stage, plain and reduction kernels.
 *
 * HBM layout of one volume subdomain (K elements, order N, Np nodes, NF
 * fields): element-major, field-major inside the element,
 *     q[e][f][n]   (state, two ping-pong copies for the fused stage)
 *     dq[e][f][n]  (LSRK rate register)
 * so a workgroup streams one contiguous NF*Np block per element and every
 * field row is a fully coalesced run of Np values.  Per-element geometry and
 * material live in small AoS tables (GEO_STRIDE / MAT_STRIDE entries), the
 * face connectivity in nbr_elem/nbr_code (one entry per element face).
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bfamx
{

/* Pointers read out of a descriptor in memory are generic (flat) to the
 * compiler; casting them into the global address space lets loads and stores
 * use global_* instructions with a scalar base and a 32-bit lane offset
 * instead of 64-bit flat addresses per lane. */
template <class T>
using gptr = __attribute__((address_space(1))) T *;

template <class T> __device__ __forceinline__ gptr<T> glob(T *p)
{
  return (gptr<T>)(p);
}

/* Read-only tables that no kernel which reads them writes (task lists,
 * subdomain descriptors, clocks inside a stage launch, operators, element
 * material and connectivity) are read through the constant address space:
 * a wave-uniform load from it becomes a scalar load through the scalar data
 * cache, instead of a vector load queued behind the kernel's HBM streams
 * whose result must then be moved to a scalar register. */
template <class T>
using cptr = const __attribute__((address_space(4))) T *;

template <class T> __device__ __forceinline__ cptr<T> cglob(const T *p)
{
  return (cptr<T>)(p);
}

/* pointer to an element table: constant (scalar loads when the element is
 * uniform in the wave) or global address space */
template <bool SCALAR, class T> struct ElemPtr
{
  using type = gptr<const T>;
};
template <class T> struct ElemPtr<true, T>
{
  using type = cptr<T>;
};

enum
{
  GEO_GX = 0, /* 2/h_x, 2/h_y, 2/h_z : reference -> physical derivative */
  GEO_GY = 1,
  GEO_GZ = 2,
  GEO_LX = 3, /* lift factors (sJ/J)/w_end = (2/h_d)/w_0 */
  GEO_LY = 4,
  GEO_LZ = 5,
  GEO_X0 = 6, /* element lower corner */
  GEO_Y0 = 7,
  GEO_Z0 = 8,
  GEO_HX = 9,
  GEO_HY = 10,
  GEO_HZ = 11,
  GEO_STRIDE = 12
};

enum
{
  MAT_RHOI = 0, /* 1/rho */
  MAT_LAM = 1,  /* lambda (acoustic: bulk modulus) */
  MAT_MU = 2,
  MAT_ZP = 3, /* rho cp */
  MAT_ZS = 4, /* rho cs */
  MAT_STRIDE = 8
};

enum
{
  FMAT_ZP = 0,   /* plus-side rho cp */
  FMAT_ZS = 1,   /* plus-side rho cs */
  FMAT_IZP = 2,  /* 1 / (Zp- + Zp+) */
  FMAT_IZS = 3,  /* 1 / (Zs- + Zs+), 0 when both sides are fluid */
  FMAT_STRIDE = 4
};

enum
{
  MODE_STAGE = 0, /* RHS + LSRK update, clock advanced by the last block */
  MODE_RHS = 1    /* write L(q, t) as FP64 into rhs, no update */
};

/* device-resident clock of one subdomain; all dynamic state lives here so a
 * captured graph can be replayed */
struct Clock
{
  double t;            /* time at the start of the current step */
  double dt;           /* step size */
  long long stage;     /* stages completed since set_fields */
  unsigned int arrive; /* blocks of the current launch that finished */
  unsigned int pad;
  double g;            /* ricker(t_stage) of the stage about to run */
  double sw, cw;       /* sin, cos(omega t_stage) for the manufactured forcing */
};

/* everything a stage needs about one subdomain (passed by value or read
 * from a device table in batched launches) */
struct SubDev
{
  int physics;
  int N;
  int K;
  int nf;
  int nstages;
  int src_elem; /* -1 none */
  int mms;      /* 1: manufactured forcing */
  /* 1: stage s >= 1 rebuilds the previous dq from the two state buffers,
   * dq_{s-1} = (q_s - q_{s-1}) / B_{s-1}, and dq is never stored (state and
   * rates of one type, A[0] = 0, B[0..S-2] nonzero) */
  int dq_rec;
  void *q[2];   /* state ping-pong (double or float) */
  void *dq;     /* rate register (double or float); NULL with dq_rec */
  double *rhs;  /* FP64 RHS (plain path and MODE_RHS) */
  const int32_t *nbr_elem;
  const uint8_t *nbr_code;
  const double *geo;
  const float *geo_f;
  const double *mat;
  const float *mat_f;
  /* per element face: Zp+, Zs+, 1/(Zp- + Zp+), 1/(Zs- + Zs+) (0 if no S) */
  const double *fmat;
  const float *fmat_f;
  const double *Dr; /* (N+1)^2 column-major */
  const float *Dr_f;
  const double *r; /* LGL nodes */
  const double *w; /* LGL weights */
  double wi0;      /* 1 / w_0: lift weight of a face node relative to 2/h */
  const double *src_wt;
  const float *src_wt_f;
  const double *mms_tab; /* K x 3 axes x {sin, cos} x (N+1): sin/cos(k x) */
  const float *mms_tab_f;
  Clock *clock;
  /* glued faces (nbr_code & NBR_GLUED): nbr_elem is the offset of the
   * face's 6 x (N+1)^2 block in this buffer (arithmetic type), written by
   * the mortar kernel before each stage: dv[3], dT[3] per face node */
  void *glue_buf;
  double A[8], B[8], C[9];
  double AB[8]; /* the stage kernels' A: A[s] / B[s-1] with dq_rec, else A[s] */
  double src_amp[9];
  double src_f0, src_t0;
  double mms_k, mms_omega;
  double mms_F[9];
};

/* nbr_code flag of a face whose flux comes from the mortars */
constexpr int NBR_GLUED = 0x80;

/* nbr_code of a physical boundary face: NBR_BOUNDARY | kind << 3 | own face.
 * The plus state is the face's own trace, mirrored by kind:
 *   1 free surface (traction-free / pressure-free): v+ = v-, sigma+ = -sigma-
 *   2 rigid wall (no-slip / no normal flow):          v+ = -v-, sigma+ = sigma-
 *   3 absorbing (characteristic):                    v+ = 0,  sigma+ = 0
 * so that the exact Riemann solution at the face meets the condition. */
constexpr int NBR_BOUNDARY = 0x40;

/* orientation bits and mirror signs of a (non-glued) face code */
__device__ __forceinline__ int nbr_perm(int code)
{
  return (code & NBR_BOUNDARY) ? 0 : (code >> 3) & 7;
}
template <class T>
__device__ __forceinline__ void nbr_mirror(int code, T *sv, T *ss)
{
  if (!(code & NBR_BOUNDARY))
  {
    *sv = T(1);
    *ss = T(1);
    return;
  }
  const int kind = (code >> 3) & 7;
  *sv = kind == 1 ? T(1) : (kind == 2 ? T(-1) : T(0));
  *ss = kind == 1 ? T(-1) : (kind == 2 ? T(1) : T(0));
}

/* glued faces are at most order 7: (N+1)^2 <= 64 nodes, one per lane */
constexpr int GLUE_MAXFP = 64;

/* one interface handled by one wavefront of the mortar kernel: the b face
 * and its nm mortars (1: conforming, orders differ, quad0 = -1; 4: hanging,
 * mortar m is quadrant m of the b face).  *_off: offsets of the faces'
 * 6 x (N+1)^2 blocks in the glue buffer */
struct MortarGroup
{
  int32_t b_sub, b_elem, b_off;
  int32_t s_sub[4], s_elem[4], s_off[4];
  int8_t b_face, nm, quad0, pad;
  int8_t s_face[4];
  int8_t s_ng[4]; /* glue grid order of each mortar; 0: max(N_s, N_b) */
};

struct GlueArgs
{
  const SubDev *subs; /* the batch's descriptors */
  const MortarGroup *groups;
  int ngroups;
  const void *ops; /* glue_ops8_index table, arithmetic type */
  void *buf;       /* face fluctuation blocks */
};

/* 1-D mortar operators from order Na to Nb (1..8), column-major
 * (Nb+1) x (Na+1) in blocks of 81: kind 0 = prj[k], kind 1 = wi_mass_prj[k],
 * k = 0..4 as in create_interpolators */
__host__ __device__ constexpr int glue_ops_index(int kind, int Na, int Nb, int k)
{
  return ((((Na - 1) * 8 + (Nb - 1)) * 2 + kind) * 5 + k) * 81;
}
constexpr int GLUE_OPS_SIZE = 64 * 2 * 5 * 81;

/* the operators between orders 1..7 (the orders a face of the mortar
 * kernel can have), each padded to 8 x 8: entry (row, col) at row + 8 col,
 * zero outside the actual size; the mortar kernel's table */
__host__ __device__ constexpr int glue_ops8_index(int kind, int Na, int Nb, int k)
{
  return ((((Na - 1) * 7 + (Nb - 1)) * 2 + kind) * 5 + k) * 64;
}
constexpr int GLUE_OPS8_SIZE = 49 * 2 * 5 * 64;

/* one work item of a stage launch: elements [e0, e0+ne) of subdomain sub */
struct Task
{
  int sub;
  int e0;
  int ne;
  int N;
};

/* a value known to be the same in every lane, moved to a scalar register */
__device__ __forceinline__ float uniform(float v)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double uniform(double v)
{
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

/* field-wise load of a task from the global task table */
__device__ __forceinline__ Task load_task(const Task *tasks, int i)
{
  cptr<Task> t = cglob(tasks) + i;
  Task r;
  r.sub = t->sub;
  r.e0 = t->e0;
  r.ne = t->ne;
  r.N = t->N;
  return r;
}

struct LaunchArgs
{
  const SubDev *subs;       /* device table (batched) */
  const Task *tasks;        /* device table */
  int ntasks;
  int mode;
  int xcd_remap;
  int nxcd;
  /* node-parallel stage: one block per task even where the one-element
   * kernel takes one element per block (an order launched beside others) */
  int whole_tasks;
  double t_rhs;  /* time for MODE_RHS */
  double g_rhs;  /* ricker(t_rhs) */
  double sw_rhs; /* sin(omega t_rhs) */
  double cw_rhs; /* cos(omega t_rhs) */
};

/* element-major offset of (element, field, node) */
__device__ __forceinline__ size_t qidx(int e, int f, int n, int NF, int Np)
{
  return ((size_t)e * NF + f) * Np + n;
}

/*
 * Order in which a block walks the task list.  Hardware dispatches
 * consecutive workgroups round-robin over the XCDs.  Mode 0 keeps that
 * order; mode 1 gives XCD x the contiguous slice [x*C, (x+1)*C) of the
 * (Morton-ordered) task list; mode 2 (the default) keeps one round-robin
 * front over the list but hands each XCD whole 2x2x2 Morton cubes.
 */
__device__ __forceinline__ int first_task(const LaunchArgs &a, int ntasks, int *stride,
                                          int *end)
{
  const int b = blockIdx.x, G = gridDim.x;
  if (a.xcd_remap == 2 && a.nxcd == 8 && G >= ntasks)
  {
    /* round-robin front, but each XCD's 8 blocks of a group of 64 take one
     * 2x2x2 Morton cube of tasks, whose inner faces then share an L2; a
     * partial last group keeps the identity */
    *stride = G;
    *end = ntasks;
    const int g0 = b & ~63;
    if (g0 + 64 > ntasks)
      return b;
    return g0 + 8 * (b & 7) + ((b & 63) >> 3);
  }
  if (a.xcd_remap != 1 || G % a.nxcd)
  {
    *stride = G;
    *end = ntasks;
    return b;
  }
  const int x = b % a.nxcd, s = b / a.nxcd, per = G / a.nxcd;
  const int C = (ntasks + a.nxcd - 1) / a.nxcd;
  const int lo = x * C;
  int hi = lo + C;
  if (hi > ntasks)
    hi = ntasks;
  *stride = per;
  *end = hi;
  return lo + s;
}

__device__ __forceinline__ int first_task(const LaunchArgs &a, int *stride, int *end)
{
  return first_task(a, a.ntasks, stride, end);
}

/* Ricker wavelet (1 - 2a) exp(-a), a = (pi f0 (t - t0))^2 */
__device__ __forceinline__ double ricker(double f0, double t0, double t)
{
  const double x = 3.14159265358979323846 * f0 * (t - t0);
  const double a = x * x;
  return (1 - 2 * a) * exp(-a);
}

/* time-dependent source/forcing scalars of the stage that starts next.  Kept
 * out of line: only the clock-advancing thread runs it, and inlining the FP64
 * exp/sincos would raise the register budget of the whole stage kernel. */
__device__ __noinline__ void clock_scalars(const SubDev &sd, Clock *c)
{
  const long long st = c->stage;
  const double ts = c->t + sd.C[(int)(st % sd.nstages)] * c->dt;
  if (sd.src_elem >= 0)
    c->g = ricker(sd.src_f0, sd.src_t0, ts);
  if (sd.mms)
  {
    double sv, cv;
    sincos(sd.mms_omega * ts, &sv, &cv);
    c->sw = sv;
    c->cw = cv;
  }
}

/* advance a subdomain's clock by one stage (run by one thread, in a kernel
 * launched after the stage kernel, so no block of the stage reads the clock
 * while it changes) */
__device__ __forceinline__ void clock_advance(const SubDev &sd)
{
  gptr<Clock> c = glob(sd.clock);
  const long long st = c->stage + 1;
  c->stage = st;
  if (st % sd.nstages == 0)
    c->t = c->t + c->dt;
  clock_scalars(sd, (Clock *)c);
}

/* stage parameters of the subdomain read at the start of a task */
struct StageCtx
{
  int parity;
  /* the dq this stage writes is never read: the last stage of a step of a
   * scheme whose first stage has A = 0 (every 2N-storage scheme here).
   * Where dq is rebuilt from the states (rec), no stage stores it */
  int dq_dead;
  double a, b, dt, t;
  double g, sw, cw;
  /* B[s-1] times the old dq is q[parity] - q[parity ^ 1], and a holds
   * A[s] / B[s-1] */
  int rec;
};

/* SD: SubDev, possibly qualified with the constant address space.  The
 * clock is constant within a stage launch: only clock_advance_kernel, a
 * separate launch, writes it. */
template <class SD>
__device__ __forceinline__ StageCtx stage_ctx(const SD &sd, const LaunchArgs &a)
{
  StageCtx c;
  const int mode = a.mode;
  const double t_rhs = a.t_rhs;
  cptr<Clock> clk = cglob((const Clock *)sd.clock);
  if (mode == MODE_RHS)
  {
    c.parity = (int)(clk->stage & 1);
    c.dq_dead = 0;
    c.rec = 0;
    c.a = 0;
    c.b = 0;
    c.dt = 0;
    c.t = t_rhs;
    c.g = a.g_rhs;
    c.sw = a.sw_rhs;
    c.cw = a.cw_rhs;
    return c;
  }
  c.g = clk->g;
  c.sw = clk->sw;
  c.cw = clk->cw;
  const long long st = clk->stage;
  const int s = (int)(st % sd.nstages);
  c.parity = (int)(st & 1);
  c.rec = sd.dq_rec;
  c.dq_dead = s == sd.nstages - 1 && sd.A[0] == 0.0;
  c.a = sd.AB[s];
  c.b = sd.B[s];
  c.dt = clk->dt;
  c.t = clk->t + sd.C[s] * c.dt;
  return c;
}


/* the neighbour's node facing node nm of face f (non-glued faces): the
 * upload admits only axis-aligned connectivity, a neighbour meeting face f
 * with its opposite face f ^ 1 and unpermuted nodes, and boundary faces,
 * whose plus trace is the face's own */
__device__ __forceinline__ int nbr_node(int code, int f, int n1, int nm)
{
  if (code & NBR_BOUNDARY)
    return nm;
  const int stride = f < 2 ? 1 : (f < 4 ? n1 : n1 * n1);
  return (f & 1) ? nm - (n1 - 1) * stride : nm + (n1 - 1) * stride;
}

/* face permutation (swap / reverse bits), same as bfamx_perm_apply */
__device__ __forceinline__ int perm_apply(int code, int n1, int a, int b)
{
  int x = (code & 4) ? b : a;
  int y = (code & 4) ? a : b;
  if (code & 1)
    x = n1 - 1 - x;
  if (code & 2)
    y = n1 - 1 - y;
  return x + y * n1;
}

/* volume node of face f, in-face indices (a, b) (bfam.c:6599-6781 order) */
template <int DIM>
__device__ __forceinline__ int face_node(int f, int n1, int a, int b)
{
  const int fixed = (f & 1) ? n1 - 1 : 0;
  if (DIM == 2)
    return (f < 2) ? fixed + n1 * a : a + n1 * fixed;
  switch (f >> 1)
  {
  case 0: return fixed + n1 * (a + n1 * b);
  case 1: return a + n1 * (fixed + n1 * b);
  default: return a + n1 * (b + n1 * fixed);
  }
}

} // namespace bfamx
