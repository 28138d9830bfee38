/*
 * bfamx.h -- C ABI of the MI355X-native DG right-hand side and low-storage
 * Runge-Kutta hot path that replaces bfam's device layer.
 *
 * Everything crossing this boundary is plain C: int status codes, plain
 * pointers and sizes, caller-owned host arrays.  Device buffers are owned by
 * the handles.  One handle set per host thread; every device call of a
 * bfamx_device_t is issued on that device's single HIP stream
 * (bfamx_device_stream()).
 *
 * Reference interfaces replaced (src/bfam.h and src/bfam.c of bfam):
 *   bfamx_dgx_t              data of bfam_subdomain_dgx_t        bfam.h:1364-1420
 *   bfamx_glue_t             bfam_subdomain_dgx_glue_data_t       bfam.h:1319-1362
 *   bfamx_lsrk_t / _init     bfamo_lsrk_t, bfamo_lsrk_init        bfam.h:1840-1863,
 *                                                                 bfam.c:8996-9090
 *   bfamx_device_create      bfamo_device_t + bfamo_device_malloc bfam.h:1740-1744,
 *                            (mem_fudge budget)                   bfam.c:8561-8589
 *   bfamx_subdomain_upload   bfamo_add_operator_* uploads and     bfam.c:8712-8828,
 *                            field allocation                     bfam.c:6574-6597
 *   bfamx_rhs/_stage/_step   the application RHS kernels launched through
 *                            BFAMO_KERNEL_RUN                     bfam.h:1673-1719
 *   (on one GPU the glue trace exchange bfamo_communicator_post_* and _wait,
 *    bfam.c:8900-8929, becomes a device-side gather inside the stage kernel)
 *   bfamx_timer_*            BFAMO_KERNEL_RUN's KERNEL_TIME path  bfam.h:1673-1708
 *   host setup mirrors       bfam_jacobi_* (bfam.c:2681-3135), gmask/buildmaps
 *                            (bfam.c:6599-6877), create_interpolators
 *                            (bfam.c:5606-5691), bfam_pcg32_* (bfam.c:8328-8387)
 *   bfamx_octree_*           p4est forest + split_dgx_subdomains + the face
 *                            mapping of glue construction        bfam.c:4684-4796,
 *                                                                 5083-5165
 *   bfamx_glue_upload        bfam_subdomain_dgx_glue_init         bfam.c:7328-7520
 *   bfamx_batch_*            one launch per stage over many subdomains: the
 *                            stage loop bfamo_lsrk_t implies over the
 *                            subdomains bfamo_lsrk_init collects bfam.c:9005-9010
 * No counterpart in the reference (the kernels they serve are absent there):
 * the FP64 reductions (energy, L2, manufactured-solution error, CFL time
 * step), the Ricker source and manufactured forcing, the mixed-order launch
 * autotuner and launch switches, and the request serving loop.
 */
#ifndef BFAMX_H
#define BFAMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* status codes (the reference aborts; bfamx returns, see abort_on_error)  */
/* ---------------------------------------------------------------------- */
#define BFAMX_OK 0
#define BFAMX_EINVAL 1       /* invalid argument or inconsistent input */
#define BFAMX_EBUDGET 2      /* HBM budget exceeded (bfam.c:8571-8578) */
#define BFAMX_EHIP 3         /* HIP runtime error */
#define BFAMX_ESTATE 4       /* object in the wrong state for the call */
#define BFAMX_EUNSUPPORTED 5 /* order / physics / precision not built */

/* message of the last error on the calling thread */
const char *bfamx_last_error(void);
/* 1: abort() on any error, as BFAM_ABORT does (bfam.h:36-38); default 0 */
void bfamx_set_abort_on_error(int on);

/* ---------------------------------------------------------------------- */
/* LSRK coefficients -- bfamo_lsrk_t / bfamo_lsrk_init (bfam.c:8996-9090)  */
/* ---------------------------------------------------------------------- */
typedef enum
{
  BFAMX_LSRK_KC54 = 0, /* same order as bfamo_lsrk_method_t, bfam.h:1848 */
  BFAMX_LSRK_FE = 1,
  BFAMX_LSRK_HEUN = 2,
  BFAMX_LSRK_W33 = 3,
  BFAMX_LSRK_NOOP = 4 /* not a scheme: falls back to KC54 like bfam.c:9014 */
} bfamx_lsrk_method_t;

typedef struct
{
  int n_stages;
  double A[8];
  double B[8];
  double C[9];
} bfamx_lsrk_t;

/* fills the FP64 tables bit-for-bit as bfamo_lsrk_init with REAL_DOUBLE */
int bfamx_lsrk_init(bfamx_lsrk_t *ts, bfamx_lsrk_method_t method);

/* ---------------------------------------------------------------------- */
/* host setup: operators and maps (replaces the bfam setup the hot path     */
/* consumes; results are checked against the reference restatement)       */
/* ---------------------------------------------------------------------- */
/* Legendre-Gauss-Lobatto nodes and weights, N >= 1 (bfam.c:3013) */
int bfamx_lgl(int N, double *r, double *w);
/* 1-D differentiation matrix, column-major Dr[i + j*(N+1)] = l_j'(r_i)
 * (bfam.c:3082, stored as Dr_N by generic_init bfam.c:7087-7155) */
int bfamx_diff_matrix(int N, const double *r, double *Dr);
/* Lagrange basis of the LGL nodes r (N+1) evaluated at x: out[j] = l_j(x) */
int bfamx_lagrange_eval(int N, const double *r, double x, double *out);
/*
 * 1-D mortar operators from order Na to Nb (create_interpolators,
 * bfam.c:5606-5691).  prj, mass_prj, wi_mass_prj each receive 5 column-major
 * blocks of (Nb+1)x(Na+1): same, coarsen-from-bottom, coarsen-from-top,
 * refine-to-bottom, refine-to-top.  *prj0_identity is 1 when Na == Nb, and
 * block 0 of prj is then left as the identity.
 */
int bfamx_mortar_ops(int Na, int Nb, double *prj, double *mass_prj,
                     double *wi_mass_prj, int *prj0_identity);

/* face node lists: gmask[f*Nfp + n], faces -x,+x,-y,+y(,-z,+z) (bfam.c:6599) */
int bfamx_face_mask(int dim, int N, int32_t *gmask);
/*
 * vmapM/vmapP from element-to-element connectivity (bfam.c:6783-6877).
 * EToF[k*Nfaces+f] = orientation*Nfaces + neighbour face, p4est codes.
 */
int bfamx_build_vmaps(int dim, int N, int32_t K, const int32_t *EToE,
                      const int8_t *EToF, int32_t *vmapM, int32_t *vmapP);
/*
 * Compress per-node vmapP into per-face connectivity: nbr_elem[k*Nfaces+f]
 * and nbr_code[k*Nfaces+f] = perm*8 + neighbour face, where perm (0..7 in 3-D,
 * 0..1 in 2-D) is bit0 = reverse first face index, bit1 = reverse second,
 * bit2 = swap.  Returns BFAMX_EINVAL if some face's plus nodes are not one
 * whole neighbour face under one of those permutations.
 */
int bfamx_face_connectivity(int dim, int N, int32_t K, const int32_t *vmapM,
                            const int32_t *vmapP, int32_t *nbr_elem,
                            uint8_t *nbr_code);

/* PCG32 (O'Neill 2014), bfam_pcg32_* (bfam.c:8328-8387) */
typedef struct
{
  uint64_t state;
  uint64_t inc;
} bfamx_pcg32_t;
void bfamx_pcg32_seed(bfamx_pcg32_t *rng, uint64_t initstate,
                      uint64_t initseq);
uint32_t bfamx_pcg32_next(bfamx_pcg32_t *rng);
uint32_t bfamx_pcg32_bounded(bfamx_pcg32_t *rng, uint32_t bound);
/* uniform double in [0,1) from one 32-bit draw */
double bfamx_pcg32_uniform(bfamx_pcg32_t *rng);

/*
 * Uniform brick of n[0] x n[1] (x n[2]) elements on [0,L]^dim.  Elements are
 * stored in Morton (z-)order when morton != 0 (as p4est stores a uniform
 * forest), else lexicographically.  Fills EToE/EToF (K*2*dim) with p4est
 * semantics (conforming, orientation 0; non-periodic faces map to
 * themselves), lex[k] = lexicographic index of element k, and per-element
 * lower corner x0 and edge lengths h (K*dim each).
 */
int bfamx_brick(int dim, const int *n, const int *periodic, const double *L,
                int morton, int32_t *EToE, int8_t *EToF, int32_t *lex,
                double *x0, double *h);

/* ---------------------------------------------------------------------- */
/* device                                                                  */
/* ---------------------------------------------------------------------- */
typedef struct bfamx_device bfamx_device_t;

/* hbm_budget_bytes = 0: 90% of the device's memory (cf. mem_fudge) */
int bfamx_device_create(int hip_device, size_t hbm_budget_bytes,
                        bfamx_device_t **dev);
int bfamx_device_destroy(bfamx_device_t *dev);
int bfamx_device_memory(const bfamx_device_t *dev, size_t *used,
                        size_t *budget);
int bfamx_device_synchronize(bfamx_device_t *dev);
/* the device's HIP stream, as a void* (hipStream_t) */
void *bfamx_device_stream(bfamx_device_t *dev);
/* number of compute units and XCDs the kernels are tiled for */
int bfamx_device_info(const bfamx_device_t *dev, int *num_cu, int *num_xcd);

/* ---------------------------------------------------------------------- */
/* volume subdomain + physics                                             */
/* ---------------------------------------------------------------------- */
typedef enum
{
  BFAMX_ACOUSTIC2D = 0, /* fields p, u, v */
  BFAMX_ELASTIC3D = 1   /* fields v1 v2 v3 s11 s22 s33 s12 s13 s23 */
} bfamx_physics_t;

typedef enum
{
  BFAMX_FP64 = 0,       /* fused stage, FP64 state, rates and arithmetic */
  BFAMX_FP64_PLAIN = 1, /* unfused volume / surface / update kernels, FP64 */
  BFAMX_FP32_RHS = 2,   /* fused, FP64 state, FP32 rates and arithmetic */
  BFAMX_FP32 = 3,       /* fused, FP32 state, rates and arithmetic */
  /* as FP32, with the volume derivative contractions on the matrix cores
   * in split-BF16x3 (operands split into two bf16 terms, three products,
   * FP32 accumulation); elastic, node-parallel stage kernel */
  BFAMX_FP32_BF16X3 = 4
} bfamx_precision_t;

/* the hot-path data of bfam_subdomain_dgx_t (bfam.h:1364-1420) */
typedef struct
{
  int dim;
  int N;
  int Np;
  int32_t K;
  const double *r;  /* N+1 LGL nodes */
  const double *w;  /* N+1 LGL weights */
  const double *wi; /* N+1 inverse weights */
  const double *Dr; /* (N+1)^2, column-major */
  const int32_t *vmapM;
  const int32_t *vmapP;
} bfamx_dgx_t;

typedef struct
{
  bfamx_physics_t physics;
  bfamx_precision_t precision;
  const double *elem_h;  /* K*dim affine element edge lengths */
  const double *elem_x0; /* K*dim lower corners (needed for MMS forcing) */
  const double *rho;     /* K */
  const double *cp;      /* K: P-wave (acoustic: sound) speed */
  const double *cs;      /* K: S-wave speed, >= 0; 0 makes the element a
                            fluid coupled to solid neighbours by the flux
                            (ignored for acoustics) */
} bfamx_physics_desc_t;

typedef struct bfamx_subdomain bfamx_subdomain_t;

int bfamx_subdomain_upload(bfamx_device_t *dev, const bfamx_dgx_t *sub,
                           const bfamx_physics_desc_t *phys,
                           const bfamx_lsrk_t *lsrk, bfamx_subdomain_t **out);
int bfamx_subdomain_free(bfamx_subdomain_t *sub);
int bfamx_subdomain_info(const bfamx_subdomain_t *sub, int *nfields,
                         size_t *device_bytes);
/* device bytes bfamx_subdomain_upload would allocate (admission control),
 * for the LSRK tables of bfamx_lsrk_init: with these FP32 and FP64 hold two
 * states and no dq */
size_t bfamx_subdomain_bytes(int dim, int N, int32_t K, bfamx_physics_t ph,
                             bfamx_precision_t prec);
/* the same for any LSRK table (NULL: KC54): FP32 and FP64 also hold a dq
 * when the table has A[0] != 0 or a zero B[s] before the last stage */
size_t bfamx_subdomain_bytes_lsrk(int dim, int N, int32_t K, bfamx_physics_t ph,
                                  bfamx_precision_t prec, const bfamx_lsrk_t *lsrk);

/*
 * Physical boundary conditions on faces that have no neighbour (faces the
 * maps send to themselves, e.g. non-periodic brick sides): kind[k*2dim + f]
 * = 0 none, 1 free surface (traction / pressure free), 2 rigid wall
 * (no-slip velocity / no normal flow), 3 absorbing (characteristic).  The
 * plus state of the upwind flux is the own trace mirrored accordingly.  The
 * reference routes such faces through _glue_boundary subdomains
 * (bfam.c:7851-7899) and leaves the condition to the application.
 */
int bfamx_set_boundary(bfamx_subdomain_t *sub, const int8_t *kind);

/* Ricker point source: rate[f][elem, n] += weights[n]*amp[f]*ricker(t) with
 * ricker = (1-2a)exp(-a), a = (pi f0 (t-t0))^2; weights = l_n(x_s)/(J w_n) */
int bfamx_set_source(bfamx_subdomain_t *sub, int32_t elem,
                     const double *weights, const double *amp, double f0,
                     double t0);
/* manufactured forcing (elastic): rate[f] += F[f] P_f(x) tau_f(t),
 * P_f = products of sin/cos(k x_d), tau = sin(wt) for v, cos(wt) for sigma */
int bfamx_set_forcing_mms(bfamx_subdomain_t *sub, double k, double omega,
                          const double *F);

/* state: nfields host arrays of K*Np doubles (bfam field layout); fields =
 * NULL starts from the zero state.  Resets the LSRK clock to t0, step dt. */
int bfamx_set_fields(bfamx_subdomain_t *sub, const double *const *fields,
                     double t0, double dt);
int bfamx_download_fields(bfamx_subdomain_t *sub, double *const *fields);
int bfamx_get_time(bfamx_subdomain_t *sub, double *t, int64_t *stages);

/* L(q, t) of the current state into host arrays (parity checks) */
int bfamx_rhs(bfamx_subdomain_t *sub, double t, double *const *rhs);
/*
 * bfamx_rhs allocating nothing (no counterpart in the reference): for
 * callers that cannot allocate device memory, such as tight serving loops
 * or harnesses that forbid it.  bfamx_rhs stages the FP64 RHS in a device
 * buffer it allocates on first use; this variant stages it in a fixed
 * 64 MiB device array of the library (one per GPU, held until the process
 * ends), so K * nfields * Np must be at most 8,388,608 (BFAMX_EINVAL
 * otherwise).  Calls that overlap on one GPU share that array and race.
 */
int bfamx_rhs_noalloc(bfamx_subdomain_t *sub, double t, double *const *rhs);
/* one LSRK stage / nsteps full steps with the subdomain's dt */
int bfamx_stage(bfamx_subdomain_t *sub);
int bfamx_step(bfamx_subdomain_t *sub, int nsteps);

/* FP64 reductions with compensated sums */
int bfamx_reduce_energy(bfamx_subdomain_t *sub, double *energy);
/* stable step dt = cfl * min_e(min_d h_e,d / cp_e) / (N+1)^2 (FP64 device
 * reduction; the reference has no CFL rule, this one is the build's) */
int bfamx_reduce_cfl_dt(bfamx_subdomain_t *sub, double cfl, double *dt);
/* sum J w |qa - qb|^2 and sum J w |qb|^2 over all fields */
int bfamx_reduce_l2_diff(bfamx_subdomain_t *a, bfamx_subdomain_t *b,
                         double *diff2, double *norm2);
/* same against the manufactured solution with amplitudes A at the current t */
int bfamx_reduce_mms_error(bfamx_subdomain_t *sub, const double *A,
                           double *diff2, double *norm2);
/* energy norm 0.5 sum J w (rho|dv|^2 + ds : C^-1 : ds) of the difference to
 * the manufactured solution, and of the manufactured solution itself */
int bfamx_reduce_mms_energy_error(bfamx_subdomain_t *sub, const double *A,
                                  double *err_energy, double *exact_energy);

/* ---------------------------------------------------------------------- */
/* batches: many subdomains (mixed N, independent problems) advanced by one */
/* launch per LSRK stage, and HIP-graph capture of the stage loop          */
/* ---------------------------------------------------------------------- */
typedef struct bfamx_batch bfamx_batch_t;

int bfamx_batch_create(bfamx_device_t *dev, bfamx_subdomain_t *const *subs,
                       int nsubs, bfamx_batch_t **out);
int bfamx_batch_free(bfamx_batch_t *b);
int bfamx_batch_step(bfamx_batch_t *b, int nsteps);
/* capture nsteps steps of the batch into a HIP graph; launch replays it */
int bfamx_batch_capture(bfamx_batch_t *b, int nsteps);
int bfamx_batch_launch_graph(bfamx_batch_t *b);
/*
 * Autotuning on the batch's own data, for elastic batches.  For every order
 * present it times `trials` stages of each stage kernel: the line kernel
 * (one wavefront per element, lanes own node lines) and the node-parallel
 * kernel (threads own nodes) with one block per task or on a resident
 * grid, at N = 3..5 with both its elements-per-block choices (the kinds of
 * bfamx_set_kernel), each with as many resident blocks per CU as registers
 * and LDS allow and capped at 3/4 and 1/2 of that, and keeps the fastest
 * kernel, block size and cap.  With
 * mixed orders it then chooses between one launch of the order-switching
 * kernel and one launch per order.  The state and results are left
 * unchanged.  Choices are cached on the device, per precision and order and
 * per precision and set of orders, and also used by single subdomains;
 * force = 1 re-times.  *choice: 0 mixed launch, 1 per order.
 */
int bfamx_batch_autotune(bfamx_batch_t *b, int trials, int force, int *choice);
/* set the mixed-order strategy by hand: 0 mixed launch, 1 one per order */
int bfamx_batch_set_strategy(bfamx_batch_t *b, int per_order);
/* the elastic stage kernel for (precision, order) by hand: 0 line,
 * 1 node-parallel, 2 node-parallel on a resident grid (the blocks that fit
 * the CUs at once, each looping over tasks), 3 and 4 as 1 and 2 with the
 * node-parallel kernel's other elements-per-block at N = 3..5 (elsewhere
 * the same as 1 and 2), -1 back to unknown (until
 * autotuned: FP32 resident grid at N = 4 and 6, node-parallel at N = 3, 5
 * and 7; FP64 node-parallel at N = 3..5, resident grid at N = 6, 7;
 * FP32_RHS and split-BF16x3 node-parallel); get returns the kernel that
 * runs */
int bfamx_set_kernel(bfamx_device_t *dev, int precision, int N, int kind);
int bfamx_get_kernel(const bfamx_device_t *dev, int precision, int N, int *kind);
/* resident-block cap of a stage kernel (kind 0 line, 1 node-parallel, 2
 * node-parallel with its other elements-per-block, kinds 3 and 4 of
 * bfamx_set_kernel) for
 * (precision, order N = 3..7) by hand: level 0 none, 1 three quarters, 2
 * half of the blocks per CU that registers and LDS allow (a request of
 * dynamic LDS the kernel does not use).  No counterpart in the reference:
 * the tile and occupancy choice its OCCA kernels left to the caller. */
int bfamx_set_occupancy(bfamx_device_t *dev, int precision, int N, int kind, int level);
int bfamx_get_occupancy(const bfamx_device_t *dev, int precision, int N, int kind,
                        int *level);

/* RHS L(q, t) of every member into host arrays rhs[i][field] (parity
 * checks); a glued batch evaluates its mortars first */
int bfamx_batch_rhs(bfamx_batch_t *b, double t, double *const *const *rhs);


/* ---------------------------------------------------------------------- */
/* timing on the device stream (BFAMO_KERNEL_RUN's KERNEL_TIME analogue)   */
/* ---------------------------------------------------------------------- */
int bfamx_timer_start(bfamx_device_t *dev);
int bfamx_timer_stop(bfamx_device_t *dev, float *ms);

/* launch tuning, the order in which blocks take elements: 0 hardware
 * round-robin over the XCDs, 1 a contiguous element range per XCD, 2 (the
 * default, measured fastest) one round-robin front with each XCD taking
 * whole 2x2x2 Morton cubes, so that their inner faces share an L2 */
int bfamx_set_tuning(bfamx_device_t *dev, int xcd_remap);
/* elastic stage launches: 1 at most the resident wavefronts, each looping
 * over tasks and streaming the next task's state into LDS while it writes
 * the current one; 0 (default, measured faster) one wavefront per task */
int bfamx_set_persistent(bfamx_device_t *dev, int on);

/* ---------------------------------------------------------------------- */
/* adapted meshes: brick forest refined inside a sphere, 2:1 face balanced, */
/* one subdomain per LGL order, and the mortars that glue them              */
/* (replaces the p4est mesh + split_dgx_subdomains + glue construction,     */
/* bfam.c:4684-4796, 5083-5165, 7328-7520, for aligned trees)               */
/* ---------------------------------------------------------------------- */
typedef struct
{
  int n_base;            /* trees per side of the periodic unit cube */
  int levels;            /* refinement levels, 0..3 */
  double center[3];      /* leaves meeting this sphere are refined */
  double radius;
  int order_by_level[4]; /* LGL order of the leaves of each level */
} bfamx_octree_desc_t;

/* one interface piece and both its sides (the reference keeps the same as
 * two glue elements, EToEm/EToFm/EToHm and EToEp/EToHp/EToOp, bfam.h:1319):
 * face s_face of element s_elem of subdomain s_sub coincides with the
 * mortar; face b_face of (b_sub, b_elem) is the same face (quad = -1, the
 * orders differ) or the big face of a 2:1 hanging interface, the mortar
 * being its quadrant quad (bit 0: upper half along the face's first
 * tangential axis, bit 1: along the second).  order: the order N_g of the
 * mortar's glue grid, as bfam's glue_order callback sets it for user glue
 * (bfam_glue_order_t, bfam.h:1211-1212, applied at bfam.c:7815-7817); 0
 * takes the reference's default max(N_s, N_b).  Otherwise it must lie in
 * max(N_s, N_b)..7, as the reference requires N_g >= both (bfam.c:7355). */
typedef struct
{
  int32_t s_sub, s_elem, b_sub, b_elem;
  int8_t s_face, b_face, quad, order;
} bfamx_mortar_t;

typedef struct bfamx_octree bfamx_octree_t;
int bfamx_octree_build(const bfamx_octree_desc_t *desc, bfamx_octree_t **out);
int bfamx_octree_free(bfamx_octree_t *o);
/* K_level: 4 counts of leaves per level */
int bfamx_octree_info(const bfamx_octree_t *o, int32_t *K, int *nsub,
                      int32_t *nmortar, int32_t *K_level);
/* subdomain s (increasing order N): element-local EToE/EToF (K*6, glued faces
 * map to themselves), glued face flags (K*6), lower corners and edge lengths
 * (K*3), and each element's index in the whole mesh */
int bfamx_octree_sub(const bfamx_octree_t *o, int s, int *N, int32_t *K,
                     int32_t *EToE, int8_t *EToF, int8_t *glued, double *x0,
                     double *h, int32_t *gid);
/* the mortars, sorted by big side */
int bfamx_octree_mortars(const bfamx_octree_t *o, bfamx_mortar_t *m);
/* per element of the whole mesh (storage order): its subdomain and its
 * index within that subdomain (the inverse of bfamx_octree_sub's gid) */
int bfamx_octree_elements(const bfamx_octree_t *o, int32_t *sub_of,
                          int32_t *local_of);

/* ---------------------------------------------------------------------- */
/* adaptation: moving the fields to a refined / coarsened mesh between steps */
/* ---------------------------------------------------------------------- */
#define BFAMX_FLAG_SAME 0    /* BFAM_FLAG_SAME, bfam.h:239 */
#define BFAMX_FLAG_COARSEN 1 /* BFAM_FLAG_COARSEN, bfam.h:236 */
#define BFAMX_FLAG_REFINE 2  /* BFAM_FLAG_REFINE, bfam.h:237 */

/*
 * Transfer maps from a source mesh to a destination mesh of the same base
 * forest, field for field bfam_domain_pxest_transfer_maps_t (bfam.h:
 * 1122-1135).  Per destination element (mesh storage order): its adapt
 * flag, its own child id (bit d: upper half along axis d; 0 for a tree
 * root), and its source element as (subdomain, element) or, when
 * coarsened, -1 - c with c indexing the coarse arrays, which list the 8
 * source children (child id, subdomain, element) in child-id order.
 */
typedef struct
{
  uint8_t *dst_to_adapt_flags;
  int8_t *dst_to_dst_chld_id;
  int32_t *dst_to_src_subd_id;
  int32_t *dst_to_src_elem_id;
  int32_t num_dst;
  int8_t *coarse_dst_to_src_chld_id;
  int32_t *coarse_dst_to_src_subd_id;
  int32_t *coarse_dst_to_src_elem_id;
  int32_t num_coarse_dst;
} bfamx_transfer_maps_t;

/* bfam_domain_pxest_transfer_maps_init (bfam.c:5205-5387): allocates the
 * arrays; more than one level of change at an element is an error
 * (BFAMX_EINVAL; the reference aborts, bfam.c:5193-5196) */
int bfamx_transfer_maps_init(bfamx_transfer_maps_t *maps,
                             const bfamx_octree_t *dst,
                             const bfamx_octree_t *src);
/* bfam_domain_pxest_transfer_maps_free (bfam.c:5389-5399) */
int bfamx_transfer_maps_free(bfamx_transfer_maps_t *maps);

/*
 * Move the state of the source subdomains to the destination subdomains
 * through the maps, on the device, with the 1-D operators the reference
 * uploads per order pair for this (Pr, Pr_cfb, Pr_cft, Pr_r2b, Pr_r2t,
 * bfam.c:8399-8404, built by create_interpolators, bfam.c:5606): h-same
 * applies Pr(N_src -> N_dst) along each axis (a copy for equal orders),
 * h-refine Pr_r2b / Pr_r2t by the destination's child-id bit per axis,
 * h-coarsen sums the 8 children's Pr_cfb / Pr_cft images.  dst_subd and
 * dst_elem give each destination element's place (bfamx_octree_elements
 * of the destination mesh).  The sources must stand between steps at one
 * time t; the destinations get that state at t with step dt, as
 * bfamx_set_fields would set them.  Elastic, any precisions (FP64
 * arithmetic, rounded to each destination's state type).
 */
int bfamx_transfer_fields(const bfamx_transfer_maps_t *maps,
                          bfamx_subdomain_t *const *src, int nsrc,
                          bfamx_subdomain_t *const *dst, int ndst,
                          const int32_t *dst_subd, const int32_t *dst_elem,
                          double dt);

/*
 * Glue: mortars between subdomains (bfamx_mortar_t, sub indices into subs),
 * the device analogue of the glue subdomains and their operators
 * (bfam_subdomain_dgx_glue_init, bfam.c:7328-7520): integer records, the
 * 1-D interpolation / mass-projection operators of all order pairs, and one
 * flux block per glued face.  The glued faces' entries of the subdomains'
 * face tables are redirected to that block until bfamx_glue_free.  Elastic
 * fused precisions, orders 3..7.  Glued subdomains advance only through a
 * batch holding the same subdomains in the same order, with the glue set.
 */
typedef struct bfamx_glue bfamx_glue_t;
int bfamx_glue_upload(bfamx_device_t *dev, bfamx_subdomain_t *const *subs,
                      int nsubs, const bfamx_mortar_t *mortars, int32_t nmortars,
                      bfamx_glue_t **out);
int bfamx_glue_free(bfamx_glue_t *g);
int bfamx_batch_set_glue(bfamx_batch_t *b, bfamx_glue_t *g);

/*
 * The reference's glue subdomains as they are (bfam_subdomain_dgx_glue_data_t,
 * bfam.h:1319-1362, filled by bfam_subdomain_dgx_glue_init, bfam.c:7480-7492):
 * one per ordered pair of volume subdomains (minus sub_m, plus sub_p, as
 * indices into the subdomain list later given to bfamx_glue_upload), with
 * K pieces: EToEm / EToFm / EToHm on the minus side, EToEp (the piece's
 * index in the partner glue, the pair (sub_p, sub_m), which is `partner`
 * in this array; itself for a subdomain's own hanging faces), EToHp and
 * EToOp, and its order N (the glue grid's N_g, which bfam's glue_order
 * callback may have raised above both sides' orders).  The arrays are the
 * reference's own (bfam_locidx_t = int32_t).
 */
typedef struct
{
  int32_t sub_m, sub_p;
  int32_t partner;
  int32_t K;
  int32_t N; /* the glue subdomain's own order N_g (0: max of the sides) */
  const int32_t *EToEm;
  const int8_t *EToFm;
  const int8_t *EToHm;
  const int32_t *EToEp;
  const int8_t *EToHp;
  const int8_t *EToOp;
} bfamx_dgx_glue_t;

/* mortar records for bfamx_glue_upload from the reference's glue
 * subdomains: one per hanging piece (from its small side) and one per
 * conforming piece between two subdomains, with the glue's order N.
 * mortars = NULL only counts.
 * Checks that every piece and its partner name each other with matching
 * hanging numbers on opposite faces; a nonzero orientation (rotated trees)
 * is BFAMX_EUNSUPPORTED. */
int bfamx_mortars_from_glue(const bfamx_dgx_glue_t *glue, int nglue,
                            bfamx_mortar_t *mortars, int32_t *nmortars);

/* ---------------------------------------------------------------------- */
/* output                                                                  */
/* ---------------------------------------------------------------------- */
/*
 * One subdomain's nodal fields (host arrays of K*(N+1)^dim doubles, bfam
 * node order) to a VTK XML unstructured-grid file, as the reference writes
 * its per-subdomain piece (bfam_subdomain_dgx_vtk_write_vtu_piece,
 * bfam.c:6050-6572, 8234-8326) for affine elements: every element split
 * into linear VTK_PIXEL / VTK_VOXEL cells through Np_write equispaced
 * points per direction (0: the LGL nodes), Int32 connectivity and offsets,
 * the cell arrays "time", "mpirank", "subdomain_id" and "root_id" (ids[0],
 * ids[1], ids[2]: the writer's rank, subdomain->id and subdomain->uid;
 * NULL writes zeros), and one point array per field.  binary: 0 ascii,
 * 1 base64 binary, 2 base64 binary compressed with zlib in 32 KiB blocks
 * (compressor="vtkZLibDataCompressor", the reference's compress flag).
 * Version-0.1 files: 32-bit header words, so an array is at most 4 GiB.
 */
int bfamx_vtk_write(const char *path, int dim, int N, int32_t K,
                    const double *elem_x0, const double *elem_h, const double *r,
                    int nfields, const char *const *names,
                    const double *const *fields, double time, int binary,
                    int Np_write, const int32_t *ids);
/*
 * The index (.pvtu) over npieces .vtu files written by bfamx_vtk_write with
 * the same field names and encoding (binary as there), so that a viewer
 * opens all subdomains as one dataset: the PUnstructuredGrid master file of
 * bfam_vtk_write_file_pvtu (bfam.c:8066-8170), one Piece per file, with
 * the same cell arrays (time, mpirank, subdomain_id, root_id).
 * piece_paths are written as given (relative to the index's directory).
 */
int bfamx_vtk_write_index(const char *path, int npieces,
                          const char *const *piece_paths, int nfields,
                          const char *const *names, int binary);

/* ---------------------------------------------------------------------- */
/* serving: independent elastic solves admitted under an HBM budget as they */
/* arrive and advanced together by one batched launch per LSRK stage.  No   */
/* counterpart in the reference (the budget check follows                   */
/* bfamo_device_malloc, bfam.c:8561-8589).                                  */
/* ---------------------------------------------------------------------- */
typedef struct
{
  int n;           /* elements per side of the periodic unit-cube brick */
  int N;           /* LGL order, 3..7 */
  int nsteps;      /* LSRK4(5) steps */
  uint64_t seed;   /* layered material and source location */
  double arrival;  /* seconds after the start of bfamx_serve */
} bfamx_request_t;

typedef struct
{
  double arrival, admitted, finished; /* seconds after the start */
  double latency;                     /* finished - arrival */
  int64_t dof;                        /* 9 K (N+1)^3 */
  double energy;                      /* final energy (FP64, compensated) */
} bfamx_request_result_t;

typedef struct
{
  double wall_s;          /* start to last completion */
  double dof_stage_per_s; /* sum of dof * 5 * nsteps over wall_s */
  double p50_s, p99_s, max_s; /* latency quantiles (nearest rank) */
  size_t peak_bytes;      /* largest resident request memory */
  int max_active;         /* most requests advanced by one launch */
  int nlaunch;            /* batched stage launches */
  int nbatches;           /* batch rebuilds (joins / leaves) */
} bfamx_serve_stats_t;

/* a Poisson stream (seeded) of nreq requests: n ~ U{n_min..n_max},
 * N ~ U{N_min..N_max}, nsteps ~ U{steps_min..steps_max} */
int bfamx_serve_generate(uint64_t seed, int nreq, double rate_per_s,
                         int n_min, int n_max, int N_min, int N_max,
                         int steps_min, int steps_max, bfamx_request_t *reqs);
/* device bytes one request holds while resident */
size_t bfamx_request_bytes(const bfamx_request_t *rq, bfamx_precision_t prec);
/* the request's subdomain, uploaded, source attached, at rest at t = 0 */
int bfamx_request_build(bfamx_device_t *dev, const bfamx_request_t *rq,
                        bfamx_precision_t prec, bfamx_subdomain_t **out,
                        double *dt);
/* the request's inputs as bfamx_request_build sets them up, without a
 * device: per-element material (K = n^3, Morton order of bfamx_brick), the
 * source element and its (N+1)^3 weights (amplitudes 1 on the three normal
 * stresses), the Ricker f0 and t0, and dt; any output may be NULL */
int bfamx_request_inputs(const bfamx_request_t *rq, double *rho, double *cp,
                         double *cs, int32_t *src_elem, double *src_weights,
                         double *f0, double *t0, double *dt);
/* run the stream in real time (arrivals sorted), advancing at most
 * max_chunk steps between admissions; res[i] for reqs[i].  A host thread
 * prepares each arrived request's mesh, material and source while the
 * device steps, so that admission costs only the upload. */
int bfamx_serve(bfamx_device_t *dev, const bfamx_request_t *reqs, int nreq,
                bfamx_precision_t prec, size_t budget_bytes, int max_chunk,
                bfamx_request_result_t *res, bfamx_serve_stats_t *stats);
/* called on the serving thread when request `index` has finished, with its
 * subdomain at the final state (valid during the call only: e.g. download
 * its fields); the energy in res is already set */
typedef void (*bfamx_serve_done_t)(int index, bfamx_subdomain_t *sub, void *user);
/* bfamx_serve handing each finished request to done (may be NULL) */
int bfamx_serve_cb(bfamx_device_t *dev, const bfamx_request_t *reqs, int nreq,
                   bfamx_precision_t prec, size_t budget_bytes, int max_chunk,
                   bfamx_serve_done_t done, void *user,
                   bfamx_request_result_t *res, bfamx_serve_stats_t *stats);

#ifdef __cplusplus
}
#endif
#endif /* BFAMX_H */
