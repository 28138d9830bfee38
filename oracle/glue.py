"""Reference-format glue subdomains of an aligned adapted mesh.

TEST INFRASTRUCTURE ONLY (see oracle.h): restates how the reference turns a
p4est mesh into glue data, so that tests can hand the reference's arrays to
the library's glue ingest.

- Face mapping entries (bfam_domain_pxest_inter_subdomain_face_mapping,
  bfam.c:4666-4796): a face whose neighbour has the same size gives one
  entry when the neighbour lies in another subdomain, or when it is the
  small side of a hanging face (h = 0, nh = its quadrant + 1 on the big
  face); a face with four smaller neighbours gives one entry per half
  (h = quadrant + 1, nh = 0).  Orientation bo = BFAM_P8EST_ORIENTATION(f,
  nf, o) (bfam.c:3630-3646), the quadrants through BFAM_PXEST_BOHTONH and
  its inverse (bfam.c:3714-3720).
- One glue subdomain per ordered subdomain pair (s, ns), its entries sorted
  in send order to number them (i), then in receive order
  (bfam_subdomain_face_send_cmp / _recv_cmp, bfam.c:1962-2046), which gives
  EToEm / EToFm / EToHm / EToOp from the entry and EToEp = i, EToHp = nh
  (bfam_subdomain_dgx_glue_init, bfam.c:7480-7492).

The mesh comes as element boxes (lower corner and edge, periodic unit cube)
in mesh order with each element's subdomain and local index; neighbours
are found from the geometry.  Quadrant numbering on a face follows p4est:
bit 0 along the face's first tangential axis, bit 1 along the second, the
axes in increasing order.  Aligned trees only (every orientation 0).
"""
import numpy as np

from . import lib


def _orientation(f, nf, o):
    return lib().orc_p8est_orientation(f, nf, o)


def _bohtonh(bo, h):
    return lib().orc_p8est_perm_to_order(bo, h)


def _bohtonh_inv(bo, h):
    return lib().orc_p8est_perm_to_order_inv(bo, h)


def _tangents(a):
    return [d for d in range(3) if d != a]


def face_mapping(x0, h, sub_of):
    """mapping entries (dicts with s, k, f, h, ns, nk, nf, nh, o) of every
    glued face, mesh order, as bfam.c:4666-4796 lists them (rank 0, no glue
    ids)"""
    K = len(x0)
    x0 = np.asarray(x0, dtype=float)
    h = np.asarray(h, dtype=float)
    unit = h.min() / 2.0
    ix = np.round(x0 / unit).astype(np.int64)
    ih = np.round(h / unit).astype(np.int64)
    F = int(round(1.0 / unit))
    by_anchor = {}
    for k in range(K):
        by_anchor[(int(ih[k, 0]),) + tuple(int(v) for v in ix[k])] = k
    out = []
    for k in range(K):
        for f in range(6):
            a, up = f // 2, f & 1
            t1, t2 = _tangents(a)
            s = int(ih[k, 0])
            nf = f ^ 1
            bo = _orientation(f, nf, 0)
            # same size neighbour
            c = [int(v) for v in ix[k]]
            c[a] = (c[a] + (s if up else -s)) % F
            j = by_anchor.get((s,) + tuple(c))
            if j is not None:
                if sub_of[j] != sub_of[k]:
                    out.append(dict(s=sub_of[k], k=k, f=f, h=0, ns=sub_of[j], nk=j, nf=nf,
                                    nh=0, o=bo))
                continue
            # double size neighbour: this face is a quadrant of its face
            S = 2 * s
            c = [int(v) for v in ix[k]]
            c[a] = (c[a] + (s if up else -S)) % F
            p = [v - (v % S) for v in c]
            p[a] = c[a]
            j = by_anchor.get((S,) + tuple(p))
            if j is not None:
                q = int((c[t1] - p[t1]) // s) + 2 * int((c[t2] - p[t2]) // s)
                out.append(dict(s=sub_of[k], k=k, f=f, h=0, ns=sub_of[j], nk=j, nf=nf,
                                nh=_bohtonh(bo, q) + 1, o=bo))
                continue
            # four half-size neighbours (quad_to_half order: p4est quadrants)
            hs = s // 2
            for q in range(4):
                c = [int(v) for v in ix[k]]
                c[a] = (c[a] + (s if up else -hs)) % F
                c[t1] += hs * (q & 1)
                c[t2] += hs * (q >> 1)
                j = by_anchor.get((hs,) + tuple(c))
                if j is None:
                    raise ValueError(f"face {f} of element {k} is not 2:1 balanced")
                out.append(dict(s=sub_of[k], k=k, f=f, h=_bohtonh_inv(bo, q) + 1,
                                ns=sub_of[j], nk=j, nf=nf, nh=0, o=bo))
    return out


def _send_key(e):
    return (e["ns"], e["s"], e["nk"], e["nf"], e["nh"])


def _recv_key(e):
    return (e["s"], e["ns"], e["k"], e["f"], e["h"])


def glue_subdomains(x0, h, sub_of, local_of, orders=None):
    """one dict per glue subdomain (ordered pair of volume subdomains that
    share glued faces): sub_m, sub_p, the glue order N (max of the two
    sides' orders, bfam.c:7814, when `orders` gives each subdomain's order;
    else 0) and the arrays EToEm, EToFm, EToHm, EToEp, EToHp, EToOp, as
    bfam_subdomain_dgx_glue_init fills them"""
    entries = face_mapping(x0, h, sub_of)
    pairs = sorted({(e["s"], e["ns"]) for e in entries})
    out = []
    for (s, ns) in pairs:
        es = [e for e in entries if e["s"] == s and e["ns"] == ns]
        es.sort(key=_send_key)
        for i, e in enumerate(es):
            e["i"] = i
        es.sort(key=_recv_key)
        out.append({
            "sub_m": s, "sub_p": ns, "K": len(es),
            "N": max(orders[s], orders[ns]) if orders is not None else 0,
            "EToEm": np.array([local_of[e["k"]] for e in es], np.int32),
            "EToFm": np.array([e["f"] for e in es], np.int8),
            "EToHm": np.array([e["h"] for e in es], np.int8),
            "EToEp": np.array([e["i"] for e in es], np.int32),
            "EToHp": np.array([e["nh"] for e in es], np.int8),
            "EToOp": np.array([e["o"] for e in es], np.int8)})
    return out
