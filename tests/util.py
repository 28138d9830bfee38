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
                                ops=(m.r, m.w, 1 / m.w, m.Dr))


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
