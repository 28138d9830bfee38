"""Physical boundary conditions (free surface, rigid wall, absorbing) in
the oracle, pinned by exact standing waves between walls, an outgoing pulse
leaving through absorbing faces, and energy dissipation.  CPU only."""
import math

import numpy as np
import pytest

import bfam_amd as bx
import bfam_amd.workloads as W
from util import oracle_problem

RHO, CP, CS = 1.3, 2.0, 1.1
LAM, MU = RHO * (CP * CP - 2 * CS * CS), RHO * CS * CS


def box(n, N, kinds, physics=bx.ELASTIC3D):
    dim = 3 if physics == bx.ELASTIC3D else 2
    L = [1.0, 0.5, 0.5][:dim]
    mesh = W.Mesh(dim, list(n), N, L=L, periodic=[0] + [1] * (dim - 1))
    K = mesh.K
    cfg = {"mesh": mesh, "rho": np.full(K, RHO), "cp": np.full(K, CP),
           "cs": np.full(K, CS) if dim == 3 else None, "physics": physics,
           "dt": W.cfl_dt(float(mesh.h.min()), CP, N, 0.5)}
    cfg["bc"] = W.boundary_kinds(mesh, kinds)
    return cfg


def standing_elastic(cfg, t, wall):
    """P standing wave along x between walls at x = 0, 1 (rigid: v1 = 0
    there; free: sigma11 = 0 there)."""
    x = cfg["mesh"].coords()[0]
    k = math.pi
    w = k * CP
    q = np.zeros((9, x.size))
    if wall == bx.BC_RIGID:
        q[0] = np.sin(k * x) * math.cos(w * t)
        s11 = RHO * CP * np.cos(k * x) * math.sin(w * t)
    else:
        q[0] = np.cos(k * x) * math.cos(w * t)
        s11 = -RHO * CP * np.sin(k * x) * math.sin(w * t)
    q[3] = s11
    q[4] = q[5] = LAM / (LAM + 2 * MU) * s11
    return q


@pytest.mark.parametrize("wall", [bx.BC_RIGID, bx.BC_FREE])
def test_elastic_standing_wave_between_walls(wall):
    cfg = box((6, 1, 1), 6, [wall, wall, 0, 0, 0, 0])
    assert (cfg["bc"] != 0).sum() == 2
    P = oracle_problem(cfg)
    nsteps = 80
    q = P.run(standing_elastic(cfg, 0.0, wall), cfg["dt"], nsteps)
    ex = standing_elastic(cfg, nsteps * cfg["dt"], wall)
    err = np.sqrt(((q - ex) ** 2).sum() / (ex ** 2).sum())
    assert err < 1e-6, err
    # the wrong wall reflects with the wrong sign: a planted kind swap fails
    bad = dict(cfg, bc=np.where(cfg["bc"] > 0, 3 - cfg["bc"], 0).astype(np.int8))
    qb = oracle_problem(bad).run(standing_elastic(cfg, 0.0, wall), cfg["dt"], nsteps)
    assert np.sqrt(((qb - ex) ** 2).sum() / (ex ** 2).sum()) > 100 * err


@pytest.mark.parametrize("wall", [bx.BC_RIGID, bx.BC_FREE])
def test_acoustic_standing_wave_between_walls(wall):
    cfg = box((6, 2), 6, [wall, wall, 0, 0], physics=bx.ACOUSTIC2D)
    x = cfg["mesh"].coords()[0]
    k, Z = math.pi, RHO * CP
    w = k * CP

    def exact(t):
        q = np.zeros((3, x.size))
        if wall == bx.BC_RIGID:
            q[0] = np.cos(k * x) * math.cos(w * t)
            q[1] = np.sin(k * x) * math.sin(w * t) / Z
        else:
            q[0] = np.sin(k * x) * math.cos(w * t)
            q[1] = -np.cos(k * x) * math.sin(w * t) / Z
        return q

    P = oracle_problem(cfg)
    nsteps = 80
    q = P.run(exact(0.0), cfg["dt"], nsteps)
    ex = exact(nsteps * cfg["dt"])
    assert np.sqrt(((q - ex) ** 2).sum() / (ex ** 2).sum()) < 1e-6


def test_absorbing_faces_let_a_pulse_out():
    """A P pulse travelling +x leaves through the absorbing face at x = 1:
    once it has crossed, almost no energy remains (no reflection)."""
    cfg = box((8, 1, 1), 5, [bx.BC_ABSORB, bx.BC_ABSORB, 0, 0, 0, 0])
    x = cfg["mesh"].coords()[0]
    f = np.exp(-((x - 0.5) / 0.08) ** 2)
    q0 = np.zeros((9, x.size))
    q0[0] = f
    q0[3] = -RHO * CP * f
    q0[4] = q0[5] = -LAM / CP * f
    P = oracle_problem(cfg)
    E0 = P.energy(q0)
    T = 1.0 / CP  # the pulse centre reaches x = 1.5
    nsteps = int(T / cfg["dt"]) + 1
    q = P.run(q0, T / nsteps, nsteps)
    assert P.energy(q) < 1e-4 * E0
    # a rigid wall instead reflects it all
    rigid = dict(cfg, bc=np.where(cfg["bc"] > 0, bx.BC_RIGID, 0).astype(np.int8))
    qr = oracle_problem(rigid).run(q0, T / nsteps, nsteps)
    assert oracle_problem(rigid).energy(qr) > 0.9 * E0


@pytest.mark.parametrize("kind", [bx.BC_FREE, bx.BC_RIGID, bx.BC_ABSORB])
def test_boundary_energy_dissipation(kind):
    rng = np.random.default_rng(kind)
    cfg = box((3, 2, 2), 4, [kind] * 2 + [0] * 4)
    P = oracle_problem(cfg)
    q = rng.standard_normal((9, cfg["mesh"].K * cfg["mesh"].Np))
    d = P.rhs(q, 0.0)
    eps = 1e-3
    assert (P.energy(q + eps * d) - P.energy(q - eps * d)) / (4 * eps) < 0
