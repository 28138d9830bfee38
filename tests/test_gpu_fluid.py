"""Fluid elements (cs = 0) on the device against the oracle: the fluid-layer
brick (a fluid layer over a PCG32 solid, periodic, Ricker source in the
fluid) through both stage kernels in every fused precision, its energy, and
a glued octree whose finest subdomain is fluid (fluid-solid mortars).
Tolerances as in the other parity tests: FP64 1e-12, FP32 arithmetic 2e-5
(max norm relative to the largest value)."""
import numpy as np
import pytest

import bfam_amd as bx
import bfam_amd.workloads as W
from util import oracle_multi, oracle_problem, rel_max

pytestmark = pytest.mark.gpu

TOL = {bx.FP64: 1e-12, bx.FP32_RHS: 2e-5, bx.FP32: 2e-5}
PRECS = [bx.FP64, bx.FP32_RHS, bx.FP32]


@pytest.fixture(scope="module")
def dev():
    d = bx.Device(0)
    yield d
    d.close()


@pytest.fixture(params=[0, 1, 2], ids=["line", "node", "node_resident"])
def kernel(request, dev):
    for prec in PRECS:
        for N in range(3, 8):
            dev.set_kernel(prec, N, request.param)
    yield request.param
    for prec in PRECS:
        for N in range(3, 8):
            dev.set_kernel(prec, N, -1)


def fluid_state(cfg, seed):
    """a random state, shear-free isotropic stress in the fluid"""
    m = cfg["mesh"]
    q = np.random.default_rng(seed).standard_normal((9, m.K * m.Np))
    wet = np.repeat(cfg["wet"], m.Np)
    q[4, wet] = q[3, wet]
    q[5, wet] = q[3, wet]
    q[6:9, wet] = 0.0
    return q


@pytest.mark.parametrize("N", [3, 5, 7])
@pytest.mark.parametrize("prec", PRECS)
def test_fluid_layer_rhs(dev, kernel, N, prec):
    print(f"fluid layer rhs N={N} prec={prec} kernel={kernel}", flush=True)
    cfg = W.elastic3d_fluid_layer(n=4, N=N, nsteps=20)
    q0 = fluid_state(cfg, N)
    sub = W.upload(dev, cfg, prec)
    sub.set_fields(q0, 0.0, cfg["dt"])
    t = 0.4 * cfg["src"]["t0"]
    got = sub.rhs(t)
    sub.close()
    want = oracle_problem(cfg).rhs(q0, t)
    assert rel_max(got, want) < TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
def test_fluid_layer_steps_and_energy(dev, kernel, prec):
    print(f"fluid layer 6 steps prec={prec} kernel={kernel}", flush=True)
    cfg = W.elastic3d_fluid_layer(n=4, N=5, nsteps=20)
    q0 = fluid_state(cfg, 11)
    sub = W.upload(dev, cfg, prec)
    sub.set_fields(q0, 0.0, cfg["dt"])
    sub.step(6)
    got = sub.download()
    E = sub.energy()
    sub.close()
    P = oracle_problem(cfg)
    want = P.run(q0, cfg["dt"], 6)
    assert rel_max(got, want) < TOL[prec]
    # the device energy of the device state, against the oracle's energy of it
    assert abs(E - P.energy(got)) < 1e-12 * P.energy(got)
    wet = np.repeat(cfg["wet"], cfg["mesh"].Np)
    assert np.abs(got[6:9, wet]).max() == 0.0


@pytest.mark.parametrize("prec", PRECS)
def test_glued_fluid_solid_octree(dev, prec):
    """2:1 hanging mortars between a fluid subdomain (the finest level) and
    solid ones: one glued RHS and three glued steps against the oracle"""
    print(f"glued fluid-solid octree prec={prec}", flush=True)
    cfg = W.elastic3d_octree(n_base=4, levels=2, radius=0.2, orders=(5, 4, 3), seed=5)
    f = len(cfg["subs"]) - 1
    cfg["rho"][f][:] = 1.0
    cfg["cp"][f][:] = 1.5
    cfg["cs"][f][:] = 0.0
    rng = np.random.default_rng(6)
    q0 = []
    for s, sm in enumerate(cfg["subs"]):
        q = rng.standard_normal((9, sm.K * sm.Np))
        if s == f:
            q[4] = q[3]
            q[5] = q[3]
            q[6:9] = 0.0
        q0.append(q)
    subs = W.upload_multi(dev, cfg, prec)
    for s, q in zip(subs, q0):
        s.set_fields(q, 0.0, cfg["dt"])
    glue = bx.Glue(dev, subs, cfg["mortars"])
    b = bx.Batch(dev, subs)
    b.set_glue(glue)
    t = 0.4 * cfg["src"]["t0"]
    got_rhs = b.rhs(t)
    b.step(3)
    got = [s.download() for s in subs]
    b.close()
    glue.close()
    for s in subs:
        s.close()
    P = oracle_multi(cfg)
    want_rhs = P.rhs(q0, t)
    want = P.run(q0, cfg["dt"], 3)

    def rel(g, w):
        return max(np.abs(x - y).max() for x, y in zip(g, w)) / max(np.abs(y).max() for y in w)

    assert rel(got_rhs, want_rhs) < TOL[prec]
    assert rel(got, want) < TOL[prec]
    assert np.abs(got[f][6:9]).max() == 0.0
