/*
 * Adapted brick forest with 2:1 face balance, split into one subdomain per
 * LGL order, and the mortar list that glues them.
 *
 * Replaces, for a periodic unit cube of n_base^3 aligned trees, what the
 * reference obtains from p4est (p8est_refine / p8est_balance /
 * p8est_mesh_new: quad_to_quad, quad_to_face, quad_to_half) and turns into
 * volume and glue subdomains: split_dgx_subdomains groups elements by order
 * (bfam.c:5083-5165) and the face mapping enumerates conforming,
 * mixed-order and hanging faces (bfam.c:4684-4796).  Elements are leaves of
 * an octree with integer anchors on the finest grid; storage order is the
 * Morton order of the anchors, as in a p4est forest of aligned trees.
 *
 * One mortar record describes one interface piece with both its sides: the
 * "s" face coincides with the mortar, the "b" face is either the same face
 * (conforming, mixed order) or the big face of a 2:1 hanging interface, of
 * which the mortar is quadrant `quad` (bit 0: upper half along the face's
 * first tangential axis, bit 1: along the second).  The reference stores the
 * same information as two glue elements per piece (EToEm/EToFm/EToHm and
 * EToEp/EToHp/EToOp, bfam.h:1319-1362); with aligned trees every
 * orientation is 0.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "bfamx_internal.h"

#define MAXLEV 4

typedef struct
{
  int level;
  int x[3]; /* anchor on the finest grid */
} leaf_t;

struct bfamx_octree
{
  int n_base, levels, F;
  int32_t K;
  leaf_t *leaf; /* storage (Morton) order */
  int *order;   /* LGL order per element */
  int nsub;
  int sub_N[MAXLEV + 1];
  int32_t sub_K[MAXLEV + 1];
  int32_t *sub_of;   /* element -> subdomain */
  int32_t *local_of; /* element -> index within its subdomain */
  int32_t *EToE;     /* per element, in-subdomain neighbour (local), 6 */
  int8_t *EToF;
  int8_t *glued; /* 6 per element: face handled by a mortar */
  int32_t nmortar;
  bfamx_mortar_t *mortar;
  int32_t K_level[MAXLEV + 1];
};

/* ---------------------------------------------------------------- hashing */

typedef struct
{
  uint64_t *key;
  int32_t *val;
  size_t cap;
} table_t;

static uint64_t leaf_key(int level, const int *x)
{
  return ((uint64_t)(level + 1) << 60) | ((uint64_t)x[0] << 40) |
         ((uint64_t)x[1] << 20) | (uint64_t)x[2];
}

static size_t hash64(uint64_t k)
{
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return (size_t)k;
}

static int table_init(table_t *t, size_t n)
{
  t->cap = 16;
  while (t->cap < 2 * n + 16)
    t->cap <<= 1;
  t->key = calloc(t->cap, sizeof(uint64_t));
  t->val = malloc(t->cap * sizeof(int32_t));
  return t->key && t->val;
}

static void table_free(table_t *t)
{
  free(t->key);
  free(t->val);
}

static void table_put(table_t *t, uint64_t k, int32_t v)
{
  size_t i = hash64(k) & (t->cap - 1);
  while (t->key[i] && t->key[i] != k)
    i = (i + 1) & (t->cap - 1);
  t->key[i] = k;
  t->val[i] = v;
}

static int32_t table_get(const table_t *t, uint64_t k)
{
  size_t i = hash64(k) & (t->cap - 1);
  while (t->key[i])
  {
    if (t->key[i] == k)
      return t->val[i];
    i = (i + 1) & (t->cap - 1);
  }
  return -1;
}

/* leaf containing fine cell c (coordinates already wrapped) */
static int32_t find_leaf(const table_t *t, int levels, const int *c)
{
  for (int l = levels; l >= 0; --l)
  {
    const int s = 1 << (levels - l);
    const int a[3] = {c[0] & ~(s - 1), c[1] & ~(s - 1), c[2] & ~(s - 1)};
    const int32_t v = table_get(t, leaf_key(l, a));
    if (v >= 0)
      return v;
  }
  return -1;
}

static int wrap(int v, int F)
{
  return ((v % F) + F) % F;
}

static int rebuild_table(table_t *t, const leaf_t *lf, int32_t n)
{
  table_free(t);
  if (!table_init(t, (size_t)n))
    return 0;
  for (int32_t i = 0; i < n; ++i)
    table_put(t, leaf_key(lf[i].level, lf[i].x), i);
  return 1;
}

/* replace the marked leaves by their 8 children; returns the new count */
static int32_t refine_marked(leaf_t **lf, int32_t n, const char *mark, int levels)
{
  int32_t nm = 0;
  for (int32_t i = 0; i < n; ++i)
    nm += mark[i] != 0;
  if (!nm)
    return n;
  leaf_t *out = malloc(sizeof(leaf_t) * (size_t)(n + 7 * nm));
  int32_t k = 0;
  for (int32_t i = 0; i < n; ++i)
  {
    const leaf_t *p = &(*lf)[i];
    if (!mark[i])
    {
      out[k++] = *p;
      continue;
    }
    const int h = 1 << (levels - p->level - 1);
    for (int c = 0; c < 8; ++c)
    {
      leaf_t ch;
      ch.level = p->level + 1;
      for (int d = 0; d < 3; ++d)
        ch.x[d] = p->x[d] + (((c >> d) & 1) ? h : 0);
      out[k++] = ch;
    }
  }
  free(*lf);
  *lf = out;
  return k;
}

static uint64_t morton3(const int *x)
{
  uint64_t k = 0;
  for (int b = 0; b < 21; ++b)
    for (int d = 0; d < 3; ++d)
      k |= (uint64_t)((x[d] >> b) & 1) << (3 * b + d);
  return k;
}

typedef struct
{
  uint64_t key;
  int32_t i;
} mkey_t;

static int mkey_cmp(const void *a, const void *b)
{
  const uint64_t x = ((const mkey_t *)a)->key, y = ((const mkey_t *)b)->key;
  return (x > y) - (x < y);
}

static int mortar_cmp(const void *a, const void *b)
{
  const bfamx_mortar_t *x = a, *y = b;
  if (x->b_sub != y->b_sub)
    return x->b_sub < y->b_sub ? -1 : 1;
  if (x->b_elem != y->b_elem)
    return x->b_elem < y->b_elem ? -1 : 1;
  if (x->b_face != y->b_face)
    return x->b_face < y->b_face ? -1 : 1;
  return x->quad - y->quad;
}

/* tangential axes of a face normal to axis a, in increasing order */
static void tangents(int a, int *t1, int *t2)
{
  *t1 = a == 0 ? 1 : 0;
  *t2 = a == 2 ? 1 : 2;
}

int bfamx_octree_build(const bfamx_octree_desc_t *d, bfamx_octree_t **out)
{
  if (!d || !out || d->n_base < 2 || d->levels < 0 || d->levels > MAXLEV - 1 ||
      d->n_base << d->levels >= (1 << 20) || d->radius < 0)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_octree_build: bad description");
  for (int l = 0; l <= d->levels; ++l)
    if (d->order_by_level[l] < 1 || d->order_by_level[l] > 8)
      return bfamx_fail(BFAMX_EINVAL, "bfamx_octree_build: bad order for level %d", l);
  const int L = d->levels, F = d->n_base << L;
  int32_t n = d->n_base * d->n_base * d->n_base;
  leaf_t *lf = malloc(sizeof(leaf_t) * (size_t)n);
  if (!lf)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_octree_build: out of host memory");
  for (int32_t i = 0; i < n; ++i)
  {
    const int tx = i % d->n_base, ty = (i / d->n_base) % d->n_base,
              tz = i / (d->n_base * d->n_base);
    lf[i].level = 0;
    lf[i].x[0] = tx << L;
    lf[i].x[1] = ty << L;
    lf[i].x[2] = tz << L;
  }
  /* refine the leaves that meet the sphere, level by level */
  for (int l = 0; l < L; ++l)
  {
    char *mark = calloc((size_t)n, 1);
    for (int32_t i = 0; i < n; ++i)
    {
      if (lf[i].level != l)
        continue;
      const double s = (double)(1 << (L - lf[i].level)) / F;
      double dist2 = 0;
      for (int dd = 0; dd < 3; ++dd)
      {
        const double lo = (double)lf[i].x[dd] / F, hi = lo + s;
        const double c = d->center[dd];
        const double e = c < lo ? lo - c : (c > hi ? c - hi : 0);
        dist2 += e * e;
      }
      mark[i] = dist2 <= d->radius * d->radius;
    }
    n = refine_marked(&lf, n, mark, L);
    free(mark);
  }
  /* 2:1 face balance */
  table_t tab = {0, 0, 0};
  for (;;)
  {
    if (!rebuild_table(&tab, lf, n))
    {
      free(lf);
      return bfamx_fail(BFAMX_EINVAL, "bfamx_octree_build: out of host memory");
    }
    char *mark = calloc((size_t)n, 1);
    int32_t nm = 0;
    for (int32_t i = 0; i < n; ++i)
    {
      if (lf[i].level < 2)
        continue;
      const int s = 1 << (L - lf[i].level);
      for (int f = 0; f < 6; ++f)
      {
        const int a = f / 2;
        int c[3] = {lf[i].x[0], lf[i].x[1], lf[i].x[2]};
        c[a] = wrap((f & 1) ? c[a] + s : c[a] - 1, F);
        const int32_t j = find_leaf(&tab, L, c);
        if (j >= 0 && lf[j].level <= lf[i].level - 2 && !mark[j])
        {
          mark[j] = 1;
          ++nm;
        }
      }
    }
    if (!nm)
    {
      free(mark);
      break;
    }
    n = refine_marked(&lf, n, mark, L);
    free(mark);
  }
  /* storage order */
  mkey_t *mk = malloc(sizeof(mkey_t) * (size_t)n);
  for (int32_t i = 0; i < n; ++i)
  {
    mk[i].key = morton3(lf[i].x);
    mk[i].i = i;
  }
  qsort(mk, (size_t)n, sizeof(mkey_t), mkey_cmp);
  bfamx_octree_t *o = calloc(1, sizeof(*o));
  o->n_base = d->n_base;
  o->levels = L;
  o->F = F;
  o->K = n;
  o->leaf = malloc(sizeof(leaf_t) * (size_t)n);
  for (int32_t i = 0; i < n; ++i)
    o->leaf[i] = lf[mk[i].i];
  free(lf);
  free(mk);
  rebuild_table(&tab, o->leaf, n);

  /* subdomains: one per distinct order, in increasing order */
  o->order = malloc(sizeof(int) * (size_t)n);
  o->sub_of = malloc(sizeof(int32_t) * (size_t)n);
  o->local_of = malloc(sizeof(int32_t) * (size_t)n);
  int present[9] = {0};
  for (int32_t i = 0; i < n; ++i)
  {
    o->order[i] = d->order_by_level[o->leaf[i].level];
    present[o->order[i]] = 1;
    o->K_level[o->leaf[i].level]++;
  }
  int sub_of_N[9];
  for (int N = 1; N <= 8; ++N)
    if (present[N])
    {
      sub_of_N[N] = o->nsub;
      o->sub_N[o->nsub++] = N;
    }
  for (int32_t i = 0; i < n; ++i)
  {
    const int s = sub_of_N[o->order[i]];
    o->sub_of[i] = s;
    o->local_of[i] = o->sub_K[s]++;
  }

  /* faces */
  o->EToE = malloc(sizeof(int32_t) * 6 * (size_t)n);
  o->EToF = malloc(6 * (size_t)n);
  o->glued = calloc(6 * (size_t)n, 1);
  size_t mcap = 64;
  o->mortar = malloc(sizeof(bfamx_mortar_t) * mcap);
  int rc = BFAMX_OK;
  for (int32_t i = 0; i < n && !rc; ++i)
  {
    const leaf_t *p = &o->leaf[i];
    const int s = 1 << (L - p->level);
    for (int f = 0; f < 6; ++f)
    {
      const int a = f / 2, up = f & 1;
      o->EToE[6 * i + f] = o->local_of[i];
      o->EToF[6 * i + f] = (int8_t)f;
      int c[3] = {p->x[0], p->x[1], p->x[2]};
      c[a] = wrap(up ? c[a] + s : c[a] - 1, F);
      const int32_t j = find_leaf(&tab, L, c);
      if (j < 0)
      {
        rc = bfamx_fail(BFAMX_EINVAL, "octree: no leaf across face %d of %d", f, i);
        break;
      }
      const leaf_t *q = &o->leaf[j];
      if (q->level == p->level)
      {
        if (o->sub_of[j] == o->sub_of[i])
        {
          o->EToE[6 * i + f] = o->local_of[j];
          o->EToF[6 * i + f] = (int8_t)(f ^ 1);
        }
        else
        {
          o->glued[6 * i + f] = 1;
          if (i < j) /* one record per conforming pair */
          {
            if ((size_t)o->nmortar == mcap)
              o->mortar = realloc(o->mortar, sizeof(bfamx_mortar_t) * (mcap *= 2));
            bfamx_mortar_t *m = &o->mortar[o->nmortar++];
            m->s_sub = o->sub_of[i];
            m->s_elem = o->local_of[i];
            m->s_face = (int8_t)f;
            m->b_sub = o->sub_of[j];
            m->b_elem = o->local_of[j];
            m->b_face = (int8_t)(f ^ 1);
            m->quad = -1;
            m->order = 0;
          }
        }
      }
      else if (q->level == p->level - 1)
        o->glued[6 * i + f] = 1; /* small side: recorded from the big side */
      else if (q->level == p->level + 1)
      {
        o->glued[6 * i + f] = 1;
        int t1, t2;
        tangents(a, &t1, &t2);
        const int h = s / 2;
        for (int qd = 0; qd < 4; ++qd)
        {
          int cc[3] = {p->x[0], p->x[1], p->x[2]};
          cc[a] = wrap(up ? p->x[a] + s : p->x[a] - h, F);
          cc[t1] = p->x[t1] + ((qd & 1) ? h : 0);
          cc[t2] = p->x[t2] + ((qd & 2) ? h : 0);
          const int32_t k = table_get(&tab, leaf_key(p->level + 1, cc));
          if (k < 0)
          {
            rc = bfamx_fail(BFAMX_EINVAL, "octree: face %d of %d is not 2:1", f, i);
            break;
          }
          if ((size_t)o->nmortar == mcap)
            o->mortar = realloc(o->mortar, sizeof(bfamx_mortar_t) * (mcap *= 2));
          bfamx_mortar_t *m = &o->mortar[o->nmortar++];
          m->s_sub = o->sub_of[k];
          m->s_elem = o->local_of[k];
          m->s_face = (int8_t)(f ^ 1);
          m->b_sub = o->sub_of[i];
          m->b_elem = o->local_of[i];
          m->b_face = (int8_t)f;
          m->quad = (int8_t)qd;
          m->order = 0;
        }
      }
      else
      {
        rc = bfamx_fail(BFAMX_EINVAL, "octree: face %d of %d breaks 2:1 balance", f, i);
        break;
      }
      if (rc)
        break;
    }
  }
  table_free(&tab);
  if (rc)
  {
    bfamx_octree_free(o);
    return rc;
  }
  qsort(o->mortar, (size_t)o->nmortar, sizeof(bfamx_mortar_t), mortar_cmp);
  *out = o;
  return BFAMX_OK;
}

int bfamx_octree_free(bfamx_octree_t *o)
{
  if (!o)
    return BFAMX_OK;
  free(o->leaf);
  free(o->order);
  free(o->sub_of);
  free(o->local_of);
  free(o->EToE);
  free(o->EToF);
  free(o->glued);
  free(o->mortar);
  free(o);
  return BFAMX_OK;
}

int bfamx_octree_info(const bfamx_octree_t *o, int32_t *K, int *nsub,
                      int32_t *nmortar, int32_t *K_level)
{
  if (!o)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_octree_info: NULL mesh");
  if (K)
    *K = o->K;
  if (nsub)
    *nsub = o->nsub;
  if (nmortar)
    *nmortar = o->nmortar;
  if (K_level)
    for (int l = 0; l < MAXLEV; ++l)
      K_level[l] = o->K_level[l];
  return BFAMX_OK;
}

int bfamx_octree_sub(const bfamx_octree_t *o, int s, int *N, int32_t *K,
                     int32_t *EToE, int8_t *EToF, int8_t *glued, double *x0,
                     double *h, int32_t *gid)
{
  if (!o || s < 0 || s >= o->nsub)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_octree_sub: bad subdomain %d", s);
  if (N)
    *N = o->sub_N[s];
  if (K)
    *K = o->sub_K[s];
  for (int32_t i = 0; i < o->K; ++i)
  {
    if (o->sub_of[i] != s)
      continue;
    const int32_t k = o->local_of[i];
    const double hs = (double)(1 << (o->levels - o->leaf[i].level)) / o->F;
    for (int f = 0; f < 6; ++f)
    {
      if (EToE)
        EToE[6 * k + f] = o->EToE[6 * i + f];
      if (EToF)
        EToF[6 * k + f] = o->EToF[6 * i + f];
      if (glued)
        glued[6 * k + f] = o->glued[6 * i + f];
    }
    for (int d = 0; d < 3; ++d)
    {
      if (x0)
        x0[3 * k + d] = (double)o->leaf[i].x[d] / o->F;
      if (h)
        h[3 * k + d] = hs;
    }
    if (gid)
      gid[k] = i;
  }
  return BFAMX_OK;
}

int bfamx_octree_mortars(const bfamx_octree_t *o, bfamx_mortar_t *m)
{
  if (!o || (o->nmortar && !m))
    return bfamx_fail(BFAMX_EINVAL, "bfamx_octree_mortars: bad arguments");
  memcpy(m, o->mortar, sizeof(bfamx_mortar_t) * (size_t)o->nmortar);
  return BFAMX_OK;
}

/* ------------------------------------------------------------ adaptation */

int bfamx_octree_elements(const bfamx_octree_t *o, int32_t *sub_of, int32_t *local_of)
{
  if (!o)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_octree_elements: NULL mesh");
  for (int32_t i = 0; i < o->K; ++i)
  {
    if (sub_of)
      sub_of[i] = o->sub_of[i];
    if (local_of)
      local_of[i] = o->local_of[i];
  }
  return BFAMX_OK;
}

/* child id of a leaf within its parent (bit d: upper half along axis d),
 * 0 for a tree root, as p4est_quadrant_child_id numbers children */
static int child_id(const leaf_t *q, int L)
{
  if (q->level == 0)
    return 0;
  const int sh = L - q->level;
  return ((q->x[0] >> sh) & 1) | (((q->x[1] >> sh) & 1) << 1) | (((q->x[2] >> sh) & 1) << 2);
}

int bfamx_transfer_maps_free(bfamx_transfer_maps_t *m)
{
  if (!m)
    return BFAMX_OK;
  free(m->dst_to_adapt_flags);
  free(m->dst_to_dst_chld_id);
  free(m->dst_to_src_subd_id);
  free(m->dst_to_src_elem_id);
  free(m->coarse_dst_to_src_chld_id);
  free(m->coarse_dst_to_src_subd_id);
  free(m->coarse_dst_to_src_elem_id);
  memset(m, 0, sizeof(*m));
  return BFAMX_OK;
}

/*
 * Every destination leaf is a source leaf (same), a child of one (refine)
 * or the parent of eight (coarsen); anything else is more than one level
 * of change and an error, as bfam_domain_pxest_transfer_maps_init aborts
 * (quadrant_compare, bfam.c:5175-5203).  Leaves are compared on the finer
 * of the two meshes' grids, so the meshes may have different refinement
 * depths.  Coarsened destinations store -1 - (their index in the coarse
 * arrays) as source subdomain and element, and their eight children in
 * child-id order, which is their storage (Morton) order, as the reference
 * walks them (bfam.c:5323-5341).
 */
int bfamx_transfer_maps_init(bfamx_transfer_maps_t *m, const bfamx_octree_t *dst,
                             const bfamx_octree_t *src)
{
  if (!m || !dst || !src)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_transfer_maps_init: bad arguments");
  memset(m, 0, sizeof(*m));
  if (dst->n_base != src->n_base)
    return bfamx_fail(BFAMX_EINVAL,
                      "transfer: the meshes have different base forests (%d^3 and %d^3 "
                      "trees)", dst->n_base, src->n_base);
  const int Lm = dst->levels > src->levels ? dst->levels : src->levels;
  const int sd = Lm - dst->levels, ss = Lm - src->levels;
  table_t tab = {0, 0, 0};
  if (!table_init(&tab, (size_t)src->K))
    return bfamx_fail(BFAMX_EINVAL, "transfer: out of host memory");
  for (int32_t i = 0; i < src->K; ++i)
  {
    const leaf_t *q = &src->leaf[i];
    const int x[3] = {q->x[0] << ss, q->x[1] << ss, q->x[2] << ss};
    table_put(&tab, leaf_key(q->level, x), i);
  }
  const int32_t nd = dst->K;
  m->num_dst = nd;
  m->dst_to_adapt_flags = malloc((size_t)nd);
  m->dst_to_dst_chld_id = malloc((size_t)nd);
  m->dst_to_src_subd_id = malloc(sizeof(int32_t) * (size_t)nd);
  m->dst_to_src_elem_id = malloc(sizeof(int32_t) * (size_t)nd);
  int32_t *kind_src = malloc(sizeof(int32_t) * 8 * (size_t)nd);
  int rc = BFAMX_OK;
  int32_t ncoarse = 0;
  if (!m->dst_to_adapt_flags || !m->dst_to_dst_chld_id || !m->dst_to_src_subd_id ||
      !m->dst_to_src_elem_id || !kind_src)
    rc = bfamx_fail(BFAMX_EINVAL, "transfer: out of host memory");
  for (int32_t i = 0; i < nd && !rc; ++i)
  {
    const leaf_t *q = &dst->leaf[i];
    const int l = q->level;
    const int x[3] = {q->x[0] << sd, q->x[1] << sd, q->x[2] << sd};
    m->dst_to_dst_chld_id[i] = (int8_t)child_id(q, dst->levels);
    int32_t j = table_get(&tab, leaf_key(l, x));
    if (j >= 0)
    {
      m->dst_to_adapt_flags[i] = BFAMX_FLAG_SAME;
      m->dst_to_src_subd_id[i] = src->sub_of[j];
      m->dst_to_src_elem_id[i] = src->local_of[j];
      continue;
    }
    if (l >= 1)
    {
      const int s = 1 << (Lm - l + 1);
      const int p[3] = {x[0] & ~(s - 1), x[1] & ~(s - 1), x[2] & ~(s - 1)};
      j = table_get(&tab, leaf_key(l - 1, p));
      if (j >= 0)
      {
        m->dst_to_adapt_flags[i] = BFAMX_FLAG_REFINE;
        m->dst_to_src_subd_id[i] = src->sub_of[j];
        m->dst_to_src_elem_id[i] = src->local_of[j];
        continue;
      }
    }
    int found = l < Lm;
    if (found)
    {
      const int h = 1 << (Lm - l - 1);
      for (int c = 0; c < 8 && found; ++c)
      {
        const int y[3] = {x[0] + ((c & 1) ? h : 0), x[1] + ((c & 2) ? h : 0),
                          x[2] + ((c & 4) ? h : 0)};
        j = table_get(&tab, leaf_key(l + 1, y));
        found = j >= 0;
        kind_src[8 * (size_t)ncoarse + c] = j;
      }
    }
    if (!found)
    {
      rc = bfamx_fail(BFAMX_EINVAL,
                      "transfer: destination element %d (level %d) is more than one "
                      "level of refinement away from the source mesh", i, l);
      break;
    }
    m->dst_to_adapt_flags[i] = BFAMX_FLAG_COARSEN;
    m->dst_to_src_subd_id[i] = -1 - ncoarse;
    m->dst_to_src_elem_id[i] = -1 - ncoarse;
    ++ncoarse;
  }
  table_free(&tab);
  if (!rc)
  {
    m->num_coarse_dst = ncoarse;
    const size_t nc = 8 * (size_t)(ncoarse ? ncoarse : 1);
    m->coarse_dst_to_src_chld_id = malloc(nc);
    m->coarse_dst_to_src_subd_id = malloc(sizeof(int32_t) * nc);
    m->coarse_dst_to_src_elem_id = malloc(sizeof(int32_t) * nc);
    if (!m->coarse_dst_to_src_chld_id || !m->coarse_dst_to_src_subd_id ||
        !m->coarse_dst_to_src_elem_id)
      rc = bfamx_fail(BFAMX_EINVAL, "transfer: out of host memory");
    for (size_t c = 0; !rc && c < 8 * (size_t)ncoarse; ++c)
    {
      const int32_t j = kind_src[c];
      m->coarse_dst_to_src_chld_id[c] = (int8_t)child_id(&src->leaf[j], src->levels);
      m->coarse_dst_to_src_subd_id[c] = src->sub_of[j];
      m->coarse_dst_to_src_elem_id[c] = src->local_of[j];
    }
  }
  free(kind_src);
  if (rc)
    bfamx_transfer_maps_free(m);
  return rc;
}
