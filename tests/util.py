"""Shared helpers: oracle problems built from the same inputs as the device
runs, and small random workloads."""
import numpy as np

import bfam_amd as bx
import bfam_amd.workloads as W
import oracle


def lame(rho, cp, cs):
    cs = np.zeros_like(rho) if cs is None else cs
    return rho * (cp * cp - 2 * cs * cs), rho * cs * cs


def oracle_problem(cfg, mms=None):
    m = cfg["mesh"]
    lam, mu = lame(cfg["rho"], cfg["cp"], cfg.get("cs"))
    phys = oracle.ACOUSTIC2D if cfg["physics"] == bx.ACOUSTIC2D else oracle.ELASTIC3D
    if mms is None and cfg.get("F") is not None:
        mms = {"k": cfg["k"], "omega": cfg["omega"], "F": cfg["F"]}
    return oracle.OracleProblem(phys, m.N, m.vmapM, m.vmapP, m.h, cfg["rho"], lam, mu,
                                x0=m.x0, src=cfg.get("src"), mms=mms,
                                ops=(m.r, m.w, 1 / m.w, m.Dr), bc=cfg.get("bc"))


def random_elastic(n, N, seed=0, L=(1.0, 1.3, 0.8), src=True):
    """Small periodic elastic brick with random per-element material."""
    rng = np.random.default_rng(seed)
    mesh = W.Mesh(3, list(n), N, L=list(L))
    K = mesh.K
    rho = rng.uniform(1.5, 3.0, K)
    cp = rng.uniform(3.0, 6.0, K)
    cs = cp / rng.uniform(1.6, 2.0, K)
    cfg = {"mesh": mesh, "rho": rho, "cp": cp, "cs": cs, "physics": bx.ELASTIC3D,
           "dt": W.cfl_dt(float(mesh.h.min()), float(cp.max()), N, 0.5)}
    if src:
        e = int(rng.integers(K))
        rs = rng.uniform(-0.9, 0.9, 3)
        cfg["src"] = {"elem": e, "weights": W.point_source_weights(mesh, e, rs),
                      "amp": [0.1, -0.2, 0.3, 1, 1, 1, 0.2, 0, -0.1], "f0": 3.0, "t0": 0.2}
    q0 = rng.standard_normal((9, K * mesh.Np))
    return cfg, q0


def random_acoustic(n, N, seed=0, L=(1.0, 0.7)):
    rng = np.random.default_rng(seed)
    mesh = W.Mesh(2, list(n), N, L=list(L))
    K = mesh.K
    rho = rng.uniform(1.0, 2.0, K)
    cp = rng.uniform(1.0, 3.0, K)
    cfg = {"mesh": mesh, "rho": rho, "cp": cp, "cs": None, "physics": bx.ACOUSTIC2D,
           "dt": W.cfl_dt(float(mesh.h.min()), float(cp.max()), N, 0.5)}
    e = int(rng.integers(K))
    rs = rng.uniform(-0.9, 0.9, 2)
    cfg["src"] = {"elem": e, "weights": W.point_source_weights(mesh, e, rs),
                  "amp": [1.0, 0.3, -0.2], "f0": 2.0, "t0": 0.1}
    q0 = rng.standard_normal((3, K * mesh.Np))
    return cfg, q0


def rel_max(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def rel_l2(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def oracle_multi(cfg):
    """OracleMulti for an adapted-mesh config (workloads.elastic3d_octree)."""
    probs = []
    for si, sm in enumerate(cfg["subs"]):
        lam, mu = lame(cfg["rho"][si], cfg["cp"][si], cfg["cs"][si])
        src = cfg["src"] if cfg.get("src") and cfg["src"]["sub"] == si else None
        probs.append(oracle.OracleProblem(oracle.ELASTIC3D, sm.N, sm.vmapM, sm.vmapP, sm.h,
                                          cfg["rho"][si], lam, mu, x0=sm.x0, src=src,
                                          ops=(sm.r, sm.w, 1 / sm.w, sm.Dr), glued=sm.glued))
    return oracle.OracleMulti(probs, cfg["mortars"])


def two_order_brick(n=4, orders=(3, 5), material=(1.3, 2.0, 1.1), seed=None):
    """Periodic n^3 brick whose elements with x < 1/2 have order orders[0] and
    the others orders[1]: two subdomains glued by conforming mixed-order
    mortars (quad = -1).  Same dict layout as workloads.elastic3d_octree."""
    b = bx.brick(3, [n, n, n], [1, 1, 1], None, morton=False)
    K = b["K"]
    side = (b["x0"][:, 0] >= 0.5).astype(int)
    local = np.zeros(K, np.int32)
    cnt = [0, 0]
    for e in range(K):
        local[e] = cnt[side[e]]
        cnt[side[e]] += 1
    EToE = b["EToE"].reshape(K, 6)
    EToF = b["EToF"].reshape(K, 6)
    subs_d = []
    for s in range(2):
        ids = np.nonzero(side == s)[0]
        k = len(ids)
        d = {"N": orders[s], "K": k, "EToE": np.zeros(6 * k, np.int32),
             "EToF": np.zeros(6 * k, np.int8), "glued": np.zeros(6 * k, np.int8),
             "x0": b["x0"][ids].copy(), "h": b["h"][ids].copy(), "gid": ids.astype(np.int32)}
        for li, e in enumerate(ids):
            for f in range(6):
                nb = EToE[e, f]
                if side[nb] == s:
                    d["EToE"][6 * li + f] = local[nb]
                    d["EToF"][6 * li + f] = EToF[e, f]
                else:
                    d["EToE"][6 * li + f] = li
                    d["EToF"][6 * li + f] = f
                    d["glued"][6 * li + f] = 1
        subs_d.append(d)
    mort = []
    for e in range(K):
        if side[e] != 0:
            continue
        for f in range(6):
            nb = EToE[e, f]
            if side[nb] == 1:
                mort.append((0, local[e], 1, local[nb], f, EToF[e, f] % 6, -1, 0))
    m = np.array(mort, dtype=bx.MORTAR_DTYPE)
    subs = [W.SubMesh(d) for d in subs_d]
    rng = np.random.default_rng(seed) if seed is not None else None
    mats = []
    for sm in subs:
        if rng is None:
            mats.append(np.tile(np.asarray(material, float), (sm.K, 1)))
        else:
            rho = rng.uniform(1.5, 3.0, sm.K)
            cp = rng.uniform(3.0, 6.0, sm.K)
            mats.append(np.stack([rho, cp, cp / rng.uniform(1.6, 2.0, sm.K)], axis=1))
    dt = min(W.cfl_dt(float(sm.h.min()), float(mt[:, 1].max()), sm.N, 0.5)
             for sm, mt in zip(subs, mats))
    return {"subs": subs, "mortars": m, "rho": [mt[:, 0].copy() for mt in mats],
            "cp": [mt[:, 1].copy() for mt in mats], "cs": [mt[:, 2].copy() for mt in mats],
            "dt": dt, "src": None, "physics": bx.ELASTIC3D}
