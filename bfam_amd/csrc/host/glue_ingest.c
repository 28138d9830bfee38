/*
 * Ingest of the reference's glue subdomains: the per-glue-element arrays of
 * bfam_subdomain_dgx_glue_data_t (bfam.h:1319-1362), as
 * bfam_subdomain_dgx_glue_init fills them (bfam.c:7328-7520), turned into
 * the two-sided mortar records bfamx_glue_upload takes.
 *
 * The reference keeps every interface piece twice, once in the glue
 * subdomain of each ordered pair of volume subdomains (minus, plus):
 *   EToEm/EToFm/EToHm  this side's element, face and hanging number
 *                      (0: the face is the piece; q+1: the piece is
 *                      quadrant q of this, the big, face);
 *   EToEp              the same piece's index in the partner glue (the
 *                      pair (plus, minus); a subdomain's own hanging faces
 *                      pair with themselves);
 *   EToHp              the hanging number on the plus side;
 *   EToOp              the orientation of the plus face.
 * A mortar record names both sides once: it is taken from the small side
 * of a hanging piece (EToHm = 0, EToHp = q + 1) and, for a conforming piece
 * between subdomains (both numbers 0), from the glue whose minus subdomain
 * has the lower index.  Quadrants of aligned faces need no orientation
 * transform (BFAM_PXEST_BOHTONH with orientation 0 is the identity,
 * bfam.c:3714-3720); other orientations are refused, as the kernels take
 * axis-aligned geometry.
 */
#include <stdlib.h>

#include "bfamx_internal.h"

static int check_glue(const bfamx_dgx_glue_t *g, int ng, int i)
{
  const bfamx_dgx_glue_t *a = &g[i];
  if (a->K < 0 || (a->K && (!a->EToEm || !a->EToFm || !a->EToHm || !a->EToEp ||
                            !a->EToHp || !a->EToOp)))
    return bfamx_fail(BFAMX_EINVAL, "glue %d: missing arrays", i);
  if (a->partner < 0 || a->partner >= ng)
    return bfamx_fail(BFAMX_EINVAL, "glue %d: partner %d out of range", i, a->partner);
  const bfamx_dgx_glue_t *p = &g[a->partner];
  if (a->N < 0 || a->N > 7 || p->N != a->N)
    return bfamx_fail(BFAMX_EINVAL, "glue %d: order %d (partner %d): orders 0..7, the same "
                                    "on both glues of a pair", i, a->N, p->N);
  if (p->sub_m != a->sub_p || p->sub_p != a->sub_m || p->K != a->K)
    return bfamx_fail(BFAMX_EINVAL,
                      "glue %d (%d -> %d, %d pieces) and its partner %d (%d -> %d, %d "
                      "pieces) do not pair", i, a->sub_m, a->sub_p, a->K, a->partner,
                      p->sub_m, p->sub_p, p->K);
  return BFAMX_OK;
}

int bfamx_mortars_from_glue(const bfamx_dgx_glue_t *glue, int nglue,
                            bfamx_mortar_t *mortars, int32_t *nmortars)
{
  if (!glue || nglue < 0 || !nmortars)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_mortars_from_glue: bad arguments");
  for (int i = 0; i < nglue; ++i)
  {
    int rc = check_glue(glue, nglue, i);
    if (rc)
      return rc;
  }
  int32_t n = 0;
  for (int i = 0; i < nglue; ++i)
  {
    const bfamx_dgx_glue_t *a = &glue[i], *p = &glue[a->partner];
    for (int32_t k = 0; k < a->K; ++k)
    {
      const int32_t kp = a->EToEp[k];
      if (kp < 0 || kp >= p->K || p->EToEp[kp] != k)
        return bfamx_fail(BFAMX_EINVAL,
                          "glue %d piece %d: EToEp names piece %d of glue %d, which does "
                          "not name it back", i, k, kp, a->partner);
      const int hm = a->EToHm[k], hp = a->EToHp[k];
      if (hm < 0 || hm > 4 || hp < 0 || hp > 4 || (hm && hp) || p->EToHm[kp] != hp ||
          p->EToHp[kp] != hm)
        return bfamx_fail(BFAMX_EINVAL, "glue %d piece %d: hanging numbers %d / %d do not "
                                        "match the partner's", i, k, hm, hp);
      if (a->EToOp[k] != 0)
        return bfamx_fail(BFAMX_EUNSUPPORTED,
                          "glue %d piece %d: orientation %d (rotated trees) is not "
                          "supported: the kernels take axis-aligned geometry",
                          i, k, (int)a->EToOp[k]);
      const int fm = a->EToFm[k], fp = p->EToFm[kp];
      if (fm < 0 || fm > 5 || fp != (fm ^ 1))
        return bfamx_fail(BFAMX_EINVAL, "glue %d piece %d: faces %d and %d are not "
                                        "opposite", i, k, fm, fp);
      if (hm)
        continue; /* the big side: recorded from the small side */
      if (!hp && (a->sub_m > a->sub_p || (a->sub_m == a->sub_p && i > a->partner)))
        continue; /* conforming: recorded from one of the two glues */
      if (!hp && a->sub_m == a->sub_p)
        return bfamx_fail(BFAMX_EINVAL, "glue %d piece %d: a conforming piece inside one "
                                        "subdomain needs no glue", i, k);
      if (mortars)
      {
        bfamx_mortar_t *m = &mortars[n];
        m->s_sub = a->sub_m;
        m->s_elem = a->EToEm[k];
        m->s_face = (int8_t)fm;
        m->b_sub = a->sub_p;
        m->b_elem = p->EToEm[kp];
        m->b_face = (int8_t)fp;
        m->quad = (int8_t)(hp ? hp - 1 : -1);
        m->order = (int8_t)a->N;
      }
      ++n;
    }
  }
  *nmortars = n;
  return BFAMX_OK;
}
