/*
 * Face masks, volume maps and the compact face connectivity the stage kernels
 * gather through.
 *
 * Conventions that must match the reference exactly (they define the data a
 * bfam caller hands over):
 *   node n = i + (N+1) j (+ (N+1)^2 k), faces -x,+x,-y,+y,-z,+z, face node
 *   order of gmask_set (bfam.c:6599-6781); vmapM/vmapP of buildmaps with
 *   p4est's face-orientation codes (bfam.c:3630-3646, 6783-6877).
 */
#include <stdlib.h>
#include <string.h>

#include "../bfamx_internal.h"

/* p8est face-pair code and code -> permutation (bfam.c:3630-3636); these are
 * p4est's orientation conventions, so the tables are necessarily the same. */
static const int8_t k_ftof_code[6][6] = {
    {0, 1, 1, 0, 0, 1}, {2, 0, 0, 1, 1, 0}, {2, 0, 0, 1, 1, 0},
    {0, 2, 2, 0, 0, 1}, {0, 2, 2, 0, 0, 1}, {2, 0, 0, 2, 2, 0}};
static const int8_t k_code_to_perm[3][4] = {
    {1, 2, 5, 6}, {0, 3, 4, 7}, {0, 4, 3, 7}};
/* buildmaps' eight node orders (bfam.c:6836-6862) as swap/flip bits */
static const int8_t k_ns_to_bits[8] = {0, 4, 1, 5, 6, 2, 7, 3};

int bfamx_face_mask(int dim, int N, int32_t *gmask)
{
  if ((dim != 2 && dim != 3) || N < 1 || !gmask)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_face_mask: bad arguments");
  const int n1 = N + 1;
  const int nfp = dim == 2 ? n1 : n1 * n1;
  for (int f = 0; f < 2 * dim; ++f)
  {
    const int axis = f / 2, fixed = (f & 1) ? N : 0;
    for (int b = 0, n = 0; b < (dim == 3 ? n1 : 1); ++b)
      for (int a = 0; a < n1; ++a, ++n)
      {
        /* the two in-face indices run over the remaining axes in order */
        int idx[3];
        idx[axis] = fixed;
        int o0 = axis == 0 ? 1 : 0;
        int o1 = axis == 2 ? 1 : 2;
        idx[o0] = a;
        if (dim == 3)
          idx[o1] = b;
        const int node = dim == 2 ? idx[0] + n1 * idx[1]
                                  : idx[0] + n1 * (idx[1] + n1 * idx[2]);
        gmask[f * nfp + n] = node;
      }
  }
  return BFAMX_OK;
}

int bfamx_build_vmaps(int dim, int N, int32_t K, const int32_t *EToE,
                      const int8_t *EToF, int32_t *vmapM, int32_t *vmapP)
{
  if ((dim != 2 && dim != 3) || N < 1 || K < 0 || !EToE || !EToF ||
      !vmapM || !vmapP)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_build_vmaps: bad arguments");
  const int n1 = N + 1, nfaces = 2 * dim;
  const int nfp = dim == 2 ? n1 : n1 * n1;
  const int Np = dim == 2 ? n1 * n1 : n1 * n1 * n1;
  int32_t *gm = malloc(sizeof(int32_t) * nfaces * nfp);
  bfamx_face_mask(dim, N, gm);
  for (int32_t k = 0; k < K; ++k)
    for (int f = 0; f < nfaces; ++f)
    {
      const int32_t k2 = EToE[nfaces * k + f];
      const int code = EToF[nfaces * k + f];
      const int f2 = code % nfaces, o = code / nfaces;
      if (k2 < 0 || k2 >= K || f2 < 0 || o >= (dim == 2 ? 2 : 4))
      {
        free(gm);
        return bfamx_fail(BFAMX_EINVAL, "bad connectivity at (%d,%d)", k, f);
      }
      int bits;
      if (dim == 2)
        bits = o ? BFAMX_PERM_FLIP_A : 0;
      else if (k2 == k && f2 == f)
        bits = 0;
      else
        bits = k_ns_to_bits[k_code_to_perm[k_ftof_code[f][f2]][o]];
      int32_t *vm = vmapM + ((size_t)k * nfaces + f) * nfp;
      int32_t *vp = vmapP + ((size_t)k * nfaces + f) * nfp;
      for (int n = 0; n < nfp; ++n)
      {
        const int a = n % n1, b = n / n1;
        vm[n] = Np * k + gm[f * nfp + n];
        vp[n] = Np * k2 + gm[f2 * nfp + bfamx_perm_apply(bits, n1, a, b)];
      }
    }
  free(gm);
  return BFAMX_OK;
}

int bfamx_face_connectivity(int dim, int N, int32_t K, const int32_t *vmapM,
                            const int32_t *vmapP, int32_t *nbr_elem,
                            uint8_t *nbr_code)
{
  if ((dim != 2 && dim != 3) || N < 1 || K < 0 || !vmapM || !vmapP ||
      !nbr_elem || !nbr_code)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_face_connectivity: bad arguments");
  const int n1 = N + 1, nfaces = 2 * dim;
  const int nfp = dim == 2 ? n1 : n1 * n1;
  const int Np = dim == 2 ? n1 * n1 : n1 * n1 * n1;
  const int nperm = dim == 2 ? 2 : 8;
  int32_t *gm = malloc(sizeof(int32_t) * nfaces * nfp);
  bfamx_face_mask(dim, N, gm);
  int rc = BFAMX_OK;
  for (int32_t k = 0; k < K && rc == BFAMX_OK; ++k)
    for (int f = 0; f < nfaces; ++f)
    {
      const int32_t *vm = vmapM + ((size_t)k * nfaces + f) * nfp;
      const int32_t *vp = vmapP + ((size_t)k * nfaces + f) * nfp;
      for (int n = 0; n < nfp; ++n)
        if (vm[n] != Np * k + gm[f * nfp + n])
        {
          rc = bfamx_fail(BFAMX_EINVAL,
                          "vmapM of element %d face %d is not the face mask",
                          k, f);
          goto done;
        }
      const int32_t k2 = vp[0] / Np;
      if (vp[0] < 0 || k2 >= K)
      {
        rc = bfamx_fail(BFAMX_EINVAL, "vmapP out of range at (%d,%d)", k, f);
        goto done;
      }
      int found = -1;
      for (int f2 = 0; f2 < nfaces && found < 0; ++f2)
        for (int pc = 0; pc < nperm && found < 0; ++pc)
        {
          int ok = 1;
          for (int n = 0; n < nfp && ok; ++n)
            ok = vp[n] == Np * k2 + gm[f2 * nfp +
                                      bfamx_perm_apply(pc, n1, n % n1, n / n1)];
          if (ok)
            found = pc * 8 + f2;
        }
      if (found < 0)
      {
        rc = bfamx_fail(BFAMX_EINVAL,
                        "plus side of element %d face %d is not one face", k,
                        f);
        goto done;
      }
      nbr_elem[(size_t)k * nfaces + f] = k2;
      nbr_code[(size_t)k * nfaces + f] = (uint8_t)found;
    }
done:
  free(gm);
  return rc;
}

/* interleave the bits of (x, y, z) */
static uint64_t morton_key(int dim, const int *idx)
{
  uint64_t key = 0;
  for (int bit = 0; bit < 21; ++bit)
    for (int d = 0; d < dim; ++d)
      key |= (uint64_t)((idx[d] >> bit) & 1) << (bit * dim + d);
  return key;
}

typedef struct
{
  uint64_t key;
  int32_t lex;
} keyed_t;

static int keyed_cmp(const void *a, const void *b)
{
  const keyed_t *x = a, *y = b;
  return x->key < y->key ? -1 : (x->key > y->key ? 1 : 0);
}

int bfamx_brick(int dim, const int *n, const int *periodic, const double *L,
                int morton, int32_t *EToE, int8_t *EToF, int32_t *lex,
                double *x0, double *h)
{
  if ((dim != 2 && dim != 3) || !n || !periodic || !L || !EToE || !EToF)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_brick: bad arguments");
  int nn[3] = {n[0], n[1], dim == 3 ? n[2] : 1};
  for (int d = 0; d < dim; ++d)
    if (nn[d] < 1 || nn[d] > (1 << 20))
      return bfamx_fail(BFAMX_EINVAL, "bfamx_brick: n[%d]=%d", d, nn[d]);
  const int64_t K64 = (int64_t)nn[0] * nn[1] * nn[2];
  if (K64 > (1 << 26))
    return bfamx_fail(BFAMX_EINVAL, "bfamx_brick: too many elements");
  const int32_t K = (int32_t)K64;
  const int nfaces = 2 * dim;

  keyed_t *order = malloc(sizeof(keyed_t) * K);
  int32_t *pos = malloc(sizeof(int32_t) * K); /* lex -> storage index */
  for (int32_t l = 0; l < K; ++l)
  {
    const int idx[3] = {l % nn[0], (l / nn[0]) % nn[1], l / (nn[0] * nn[1])};
    order[l].key = morton ? morton_key(dim, idx) : (uint64_t)l;
    order[l].lex = l;
  }
  if (morton)
    qsort(order, K, sizeof(keyed_t), keyed_cmp);
  for (int32_t k = 0; k < K; ++k)
    pos[order[k].lex] = k;

  for (int32_t k = 0; k < K; ++k)
  {
    const int32_t l = order[k].lex;
    const int idx[3] = {l % nn[0], (l / nn[0]) % nn[1], l / (nn[0] * nn[1])};
    if (lex)
      lex[k] = l;
    for (int d = 0; d < dim; ++d)
    {
      const double hd = L[d] / nn[d];
      if (h)
        h[(size_t)k * dim + d] = hd;
      if (x0)
        x0[(size_t)k * dim + d] = idx[d] * hd;
    }
    for (int f = 0; f < nfaces; ++f)
    {
      const int d = f / 2, s = (f & 1) ? 1 : -1;
      int j[3] = {idx[0], idx[1], idx[2]};
      j[d] += s;
      int boundary = 0;
      if (j[d] < 0 || j[d] >= nn[d])
      {
        if (periodic[d])
          j[d] = (j[d] + nn[d]) % nn[d];
        else
          boundary = 1;
      }
      if (boundary)
      {
        EToE[(size_t)k * nfaces + f] = k;
        EToF[(size_t)k * nfaces + f] = (int8_t)f;
      }
      else
      {
        EToE[(size_t)k * nfaces + f] = pos[j[0] + nn[0] * (j[1] + nn[1] * j[2])];
        EToF[(size_t)k * nfaces + f] = (int8_t)(f ^ 1);
      }
    }
  }
  free(order);
  free(pos);
  return BFAMX_OK;
}
