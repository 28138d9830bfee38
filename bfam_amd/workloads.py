"""Synthetic workloads of the benchmark configurations, built with the library's
own host setup (operators, brick meshes, PCG32 materials, Ricker sources).

Config names:
  acoustic2d_plane_wave   16x16 periodic quads, N=4, plane wave, 100 KC54 steps
  elastic3d_ricker        48^3 periodic hexes, N=7, PCG32 material, Ricker
  elastic3d_octree        16^3 forest refined 2 levels inside a sphere, 2:1
                          balanced, N = 7, 5, 3 by level, mortar glue
  elastic3d_mms           32^3 periodic hexes, N=7, manufactured solution
The time step used everywhere is dt = CFL * h_min / (c_max (N+1)^2); the
reference has no CFL rule, so this one is stated here and used by the
oracle and the device runs alike.
"""
import math

import numpy as np

import bfam_amd as bx


def ops(N):
    r, w = bx.lgl(N)
    return r, w, bx.diff_matrix(N, r)


def cfl_dt(hmin, cmax, N, cfl=0.5):
    return cfl * hmin / (cmax * (N + 1) ** 2)


class Mesh:
    """Brick mesh + LGL operators + volume maps for order N."""

    def __init__(self, dim, n, N, L=None, periodic=None, morton=True):
        self.dim, self.N = dim, N
        periodic = periodic if periodic is not None else [1] * dim
        m = bx.brick(dim, n, periodic, L, morton)
        self.__dict__.update(m)
        self.r, self.w, self.Dr = ops(N)
        self.vmapM, self.vmapP = bx.build_vmaps(dim, N, self.EToE, self.EToF)
        self.Np = (N + 1) ** dim

    def coords(self):
        """x[d] of shape (K*Np,) for every volume node (bfam node order)."""
        n1 = self.N + 1
        idx = np.arange(self.Np)
        loc = [idx % n1, (idx // n1) % n1, idx // (n1 * n1)]
        out = []
        for d in range(self.dim):
            ref = 0.5 * (self.r[loc[d]] + 1.0)
            out.append((self.x0[:, d][:, None] + ref[None, :] * self.h[:, d][:, None]).ravel())
        return out

    def locate(self, x):
        """storage index of the element containing point x, and its reference
        coordinates in [-1, 1]^dim"""
        n = np.array(self.n)
        L = np.array(self.L)
        hh = L / n
        ijk = np.minimum((np.array(x) / hh).astype(int), n - 1)
        lexi = int(ijk[0] + n[0] * (ijk[1] + (n[1] * ijk[2] if self.dim == 3 else 0)))
        e = int(np.nonzero(self.lex == lexi)[0][0])
        rs = 2.0 * (np.array(x) - self.x0[e]) / self.h[e] - 1.0
        return e, rs


def pcg_material(K_lex, seed, rho=(1.5, 3.0), cp=(3.0, 6.0), ratio=(1.6, 2.0)):
    """Per-element rho ~ U[rho], cp ~ U[cp], cs = cp / U[ratio], drawn in
    lexicographic element order from PCG32(seed, 54)."""
    g = bx.PCG32(seed, 54)
    out = np.zeros((K_lex, 3))
    for l in range(K_lex):
        rr = g.uniform(*rho)
        c = g.uniform(*cp)
        out[l] = (rr, c, c / g.uniform(*ratio))
    return out


def point_source_weights(mesh, e, rs):
    """weights[n] = l_i(r) l_j(s) (l_k(t)) / (J w_i w_j (w_k)): the DG projection
    M^{-1} phi(x_s) of a point source on the LGL (diagonal) mass matrix."""
    N, n1 = mesh.N, mesh.N + 1
    lx = [bx.lagrange_eval(N, mesh.r, rs[d]) for d in range(mesh.dim)]
    J = np.prod(0.5 * mesh.h[e])
    idx = np.arange(mesh.Np)
    loc = [idx % n1, (idx // n1) % n1, idx // (n1 * n1)]
    wgt = np.ones(mesh.Np)
    for d in range(mesh.dim):
        wgt *= lx[d][loc[d]] / mesh.w[loc[d]]
    return wgt / J


def elastic3d_ricker(n=48, N=7, seed=2024, cfl=0.5, nsteps=200, L=1.0):
    """Config: 3-D isotropic elastic, uniform n^3 periodic hex brick, order N,
    PCG32 per-element material, explosive Ricker point source at a seeded
    interior point.  Ricker f0 is chosen so the wavelet peaks mid-run."""
    mesh = Mesh(3, [n, n, n], N, L=[L, L, L])
    mat_lex = pcg_material(mesh.K, seed)
    mat = mat_lex[mesh.lex]
    hmin = float(mesh.h.min())
    dt = cfl_dt(hmin, float(mat[:, 1].max()), N, cfl)
    T = nsteps * dt
    g = bx.PCG32(seed, 7)
    xs = [L * g.uniform(0.25, 0.75) for _ in range(3)]
    e, rs = mesh.locate(xs)
    src = {"elem": e, "weights": point_source_weights(mesh, e, rs),
           "amp": [0, 0, 0, 1, 1, 1, 0, 0, 0], "t0": 0.5 * T, "f0": 2.4 / T, "x": xs}
    return {"mesh": mesh, "rho": mat[:, 0].copy(), "cp": mat[:, 1].copy(),
            "cs": mat[:, 2].copy(), "dt": dt, "nsteps": nsteps, "src": src,
            "physics": bx.ELASTIC3D}


def elastic3d_fluid_layer(n=16, N=5, seed=2024, cfl=0.5, nsteps=200, depth=0.375,
                          fluid=(1.0, 1.5), L=1.0):
    """Config: elastic3d_ricker's brick with a fluid layer: the elements whose
    centre lies above z = L (1 - depth) hold a fluid (rho, cp = `fluid`, cs =
    0), the rest the PCG32 solid material.  The brick is periodic, so the
    layer meets the solid at two horizontal interfaces.  Explosive Ricker
    source at a seeded point inside the fluid layer."""
    cfg = elastic3d_ricker(n=n, N=N, seed=seed, cfl=cfl, nsteps=nsteps, L=L)
    mesh = cfg["mesh"]
    zc = mesh.x0[:, 2] + 0.5 * mesh.h[:, 2]
    wet = zc > L * (1.0 - depth)
    cfg["rho"][wet] = fluid[0]
    cfg["cp"][wet] = fluid[1]
    cfg["cs"][wet] = 0.0
    g = bx.PCG32(seed, 11)
    xs = [L * g.uniform(0.25, 0.75), L * g.uniform(0.25, 0.75),
          L * (1.0 - depth * g.uniform(0.3, 0.7))]
    e, rs = mesh.locate(xs)
    T = nsteps * cfg["dt"]
    cfg["src"] = {"elem": e, "weights": point_source_weights(mesh, e, rs),
                  "amp": [0, 0, 0, 1, 1, 1, 0, 0, 0], "t0": 0.5 * T, "f0": 2.4 / T, "x": xs}
    cfg["wet"] = wet
    return cfg


def acoustic2d_plane_wave(n=16, N=4, cfl=0.5, nsteps=100, rho=1.0, c=1.0):
    """Config: 2-D acoustic on the periodic unit square, p = cos(2 pi (x+y)),
    velocity of the plane wave travelling along (1,1)/sqrt(2)."""
    mesh = Mesh(2, [n, n], N, L=[1.0, 1.0])
    x, y = mesh.coords()
    p = np.cos(2 * np.pi * (x + y))
    u = p / (rho * c * math.sqrt(2.0))
    q = np.stack([p, u, u])
    dt = cfl_dt(1.0 / n, c, N, cfl)
    K = mesh.K
    return {"mesh": mesh, "rho": np.full(K, rho), "cp": np.full(K, c), "cs": None,
            "dt": dt, "nsteps": nsteps, "q0": q, "physics": bx.ACOUSTIC2D,
            "omega": 2 * np.pi * math.sqrt(2.0) * c}


def acoustic2d_exact(cfg, t):
    x, y = cfg["mesh"].coords()
    c = float(cfg["cp"][0])
    rho = float(cfg["rho"][0])
    p = np.cos(2 * np.pi * (x + y) - cfg["omega"] * t)
    u = p / (rho * c * math.sqrt(2.0))
    return np.stack([p, u, u])


MMS_A = np.array([1.0, 0.7, -0.4, 0.5, -0.3, 0.8, 0.35, -0.25, 0.6])


def mms_forcing_coef(A, k, omega, rho, lam, mu):
    """F = dq/dt - L q for the manufactured solution
    v_f = A_f P_f(x) cos(w t), sigma_f = A_f P_f(x) sin(w t)."""
    sv = A[0] + A[1] + A[2]
    F = np.zeros(9)
    F[0] = -omega * A[0] - k * (-A[3] + A[6] + A[7]) / rho
    F[1] = -omega * A[1] - k * (A[6] - A[4] + A[8]) / rho
    F[2] = -omega * A[2] - k * (A[7] + A[8] - A[5]) / rho
    F[3] = omega * A[3] - k * (lam * sv + 2 * mu * A[0])
    F[4] = omega * A[4] - k * (lam * sv + 2 * mu * A[1])
    F[5] = omega * A[5] - k * (lam * sv + 2 * mu * A[2])
    F[6] = omega * A[6] + mu * k * (A[0] + A[1])
    F[7] = omega * A[7] + mu * k * (A[0] + A[2])
    F[8] = omega * A[8] + mu * k * (A[1] + A[2])
    return F


def mms_exact(mesh, A, k, omega, t):
    x, y, z = mesh.coords()
    S = [np.sin(k * x), np.sin(k * y), np.sin(k * z)]
    Cc = [np.cos(k * x), np.cos(k * y), np.cos(k * z)]
    pat = [S[0] * Cc[1] * Cc[2], Cc[0] * S[1] * Cc[2], Cc[0] * Cc[1] * S[2],
           Cc[0] * Cc[1] * Cc[2], Cc[0] * Cc[1] * Cc[2], Cc[0] * Cc[1] * Cc[2],
           S[0] * S[1] * Cc[2], S[0] * Cc[1] * S[2], Cc[0] * S[1] * S[2]]
    tv, ts = math.cos(omega * t), math.sin(omega * t)
    return np.stack([A[f] * pat[f] * (tv if f < 3 else ts) for f in range(9)])


def elastic3d_mms(n=32, N=7, cfl=0.5, rho=1.0, cp=math.sqrt(3.0), cs=1.0, omega=None):
    """Config: manufactured smooth sin-product solution with forcing on the
    periodic unit cube, homogeneous material."""
    mesh = Mesh(3, [n, n, n], N, L=[1.0, 1.0, 1.0])
    k = 2 * np.pi
    omega = omega if omega is not None else 2 * np.pi
    lam = rho * (cp * cp - 2 * cs * cs)
    mu = rho * cs * cs
    F = mms_forcing_coef(MMS_A, k, omega, rho, lam, mu)
    K = mesh.K
    dt = cfl_dt(1.0 / n, cp, N, cfl)
    return {"mesh": mesh, "rho": np.full(K, rho), "cp": np.full(K, cp), "cs": np.full(K, cs),
            "dt": dt, "k": k, "omega": omega, "F": F, "A": MMS_A.copy(),
            "lam": lam, "mu": mu, "physics": bx.ELASTIC3D}


class SubMesh:
    """One subdomain of an adapted mesh: order N, its elements' geometry,
    in-subdomain maps (glued faces map to themselves) and glued flags."""

    def __init__(self, d):
        self.dim, self.N, self.K = 3, d["N"], d["K"]
        self.EToE, self.EToF, self.glued = d["EToE"], d["EToF"], d["glued"]
        self.x0, self.h, self.gid = d["x0"], d["h"], d["gid"]
        self.r, self.w, self.Dr = ops(self.N)
        self.vmapM, self.vmapP = bx.build_vmaps(3, self.N, self.EToE, self.EToF)
        self.Np = (self.N + 1) ** 3

    def coords(self):
        return Mesh.coords(self)


def elastic3d_octree(n_base=16, levels=2, seed=2024, radius=0.2, orders=(7, 5, 3),
                     cfl=0.5, nsteps=200, material="pcg", center=None):
    """Config: periodic unit cube of n_base^3 trees, leaves meeting a seeded
    sphere (or one centred at `center`) refined `levels` times, 2:1 face
    balanced; leaves of level l have order orders[l]; one subdomain per
    order, glued by mortars.  Material per element from PCG32 in global
    (Morton) element order, or homogeneous (material = (rho, cp, cs));
    Ricker source at a seeded point."""
    g = bx.PCG32(seed, 13)
    seeded = [g.uniform(0.3, 0.7) for _ in range(3)]
    center = list(center) if center is not None else seeded
    o = bx.octree(n_base, levels, center, radius, list(orders)[:levels + 1])
    subs = [SubMesh(d) for d in o["subs"]]
    if material == "pcg":
        mat = pcg_material(o["K"], seed)
    else:
        mat = np.tile(np.asarray(material, dtype=float), (o["K"], 1))
    rho = [mat[sm.gid, 0].copy() for sm in subs]
    cp = [mat[sm.gid, 1].copy() for sm in subs]
    cs = [mat[sm.gid, 2].copy() for sm in subs]
    dt = min(cfl_dt(float(sm.h.min()), float(c.max()), sm.N, cfl) for sm, c in zip(subs, cp))
    T = nsteps * dt
    xs = np.array([g.uniform(0.25, 0.75) for _ in range(3)])
    src = None
    for si, sm in enumerate(subs):
        inside = np.all((sm.x0 <= xs) & (xs < sm.x0 + sm.h), axis=1)
        if inside.any():
            e = int(np.nonzero(inside)[0][0])
            rs = 2.0 * (xs - sm.x0[e]) / sm.h[e] - 1.0
            src = {"sub": si, "elem": e, "weights": point_source_weights(sm, e, rs),
                   "amp": [0, 0, 0, 1, 1, 1, 0, 0, 0], "t0": 0.5 * T, "f0": 2.4 / T,
                   "x": xs.tolist()}
            break
    return {"octree": o, "subs": subs, "mortars": o["mortars"], "rho": rho, "cp": cp,
            "octree_args": (n_base, levels, center, radius, list(orders)[:levels + 1]),
            "cs": cs, "dt": dt, "nsteps": nsteps, "src": src, "center": center,
            "radius": radius, "K_level": o["K_level"], "physics": bx.ELASTIC3D,
            "dof": sum(9 * sm.K * sm.Np for sm in subs)}


def adaptation_pair(n_base=4, levels=1, centers=((0.3, 0.4, 0.5), (0.55, 0.4, 0.5)),
                    radius=0.15, orders=(5, 3), material=(1.3, 2.0, 1.1)):
    """Two adapted meshes of one base forest whose refined sphere moves from
    centers[0] to centers[1] (some leaves refine, some coarsen, the rest stay),
    and the transfer maps from the first to the second."""
    a = elastic3d_octree(n_base, levels, radius=radius, orders=orders, material=material,
                         center=centers[0])
    b = elastic3d_octree(n_base, levels, radius=radius, orders=orders, material=material,
                         center=centers[1])
    return a, b, bx.transfer_maps(b["octree_args"], a["octree_args"])


def boundary_kinds(mesh, kind_of_face):
    """(K, 2 dim) int8 boundary kinds: kind_of_face[f] on every face f that
    has no neighbour (maps to itself), 0 elsewhere."""
    nf = 2 * mesh.dim
    E = mesh.EToE.reshape(mesh.K, nf)
    F = mesh.EToF.reshape(mesh.K, nf)
    out = np.zeros((mesh.K, nf), np.int8)
    for f in range(nf):
        self_face = (E[:, f] == np.arange(mesh.K)) & (F[:, f] == f)
        out[self_face, f] = kind_of_face[f]
    return out


def upload(dev, cfg, precision, lsrk_method=bx.LSRK_KC54):
    """bfam_amd.Subdomain for a config dict (boundaries, source and forcing
    attached)."""
    m = cfg["mesh"]
    sub = bx.Subdomain(dev, cfg["physics"], precision, m.N, m.r, m.w, m.Dr, m.vmapM, m.vmapP,
                       m.h, cfg["rho"], cfg["cp"], cfg.get("cs"), x0=m.x0,
                       lsrk_method=lsrk_method)
    if cfg.get("bc") is not None:
        sub.set_boundary(cfg["bc"])
    if cfg.get("src") is not None:
        s = cfg["src"]
        sub.set_source(s["elem"], s["weights"], s["amp"], s["f0"], s["t0"])
    if cfg.get("F") is not None:
        sub.set_forcing_mms(cfg["k"], cfg["omega"], cfg["F"])
    return sub


def upload_multi(dev, cfg, precision, lsrk_method=bx.LSRK_KC54):
    """One bfam_amd.Subdomain per subdomain of an adapted-mesh config, with
    the source attached to the subdomain that holds it (no glue yet)."""
    subs = []
    for si, sm in enumerate(cfg["subs"]):
        sub = bx.Subdomain(dev, bx.ELASTIC3D, precision, sm.N, sm.r, sm.w, sm.Dr, sm.vmapM,
                           sm.vmapP, sm.h, cfg["rho"][si], cfg["cp"][si], cfg["cs"][si],
                           x0=sm.x0, lsrk_method=lsrk_method)
        s = cfg.get("src")
        if s is not None and s["sub"] == si:
            sub.set_source(s["elem"], s["weights"], s["amp"], s["f0"], s["t0"])
        subs.append(sub)
    return subs
