"""Serving through the C ABI: requests admitted as they arrive under an HBM
budget and advanced together by one batched launch per stage.  Each
request's result is the same alone and in the mix, bit for bit (every
element is computed by the same code whatever else shares the launch)."""
import numpy as np
import pytest

import bfam_amd as bx
import bfam_amd.workloads as W
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = bx.Device(0)
    yield d
    d.close()


def alone_energy(dev, req, prec):
    sub, dt = bx.request_build(dev, req, prec)
    sub.step(req.nsteps)
    e = sub.energy()
    t, st = sub.time()
    sub.close()
    assert st == 5 * req.nsteps and abs(t - req.nsteps * dt) < 1e-12
    return e


@pytest.mark.parametrize("prec", [bx.FP32_RHS, bx.FP32])
def test_serve_alone_equals_mix(dev, prec):
    print(f"serve: 10 requests, alone vs mix, prec={prec}", flush=True)
    reqs = bx.serve_generate(seed=3, nreq=10, rate_per_s=200.0, n=(3, 6), N=(3, 7),
                             steps=(10, 40))
    res, st = bx.serve(dev, reqs, prec, budget_bytes=1 << 30, max_chunk=7)
    print(st, flush=True)
    assert st["max_active"] >= 2  # the requests did share launches
    for q, r in zip(reqs, res):
        assert r["finished"] >= r["admitted"] >= r["arrival"] - 1e-9
        assert r["latency"] > 0 and r["energy"] > 0
        assert r["energy"] == alone_energy(dev, q, prec)
    lat = sorted(r["latency"] for r in res)
    assert st["p50_s"] == lat[4] and st["p99_s"] == lat[-1]


def test_serve_budget_limits_residency(dev):
    print("serve: budget for two at a time", flush=True)
    reqs = bx.serve_generate(seed=8, nreq=6, rate_per_s=1000.0, n=(4, 4), N=(5, 5),
                             steps=(5, 12))
    one = bx.request_bytes(reqs[0], bx.FP32)
    res, st = bx.serve(dev, reqs, bx.FP32, budget_bytes=2 * one + one // 2, max_chunk=3)
    assert st["max_active"] == 2 and st["peak_bytes"] <= 2 * one + one // 2
    for q, r in zip(reqs, res):
        assert r["energy"] == alone_energy(dev, q, bx.FP32)


def test_serve_rejects_oversized_request(dev):
    reqs = bx.serve_generate(seed=1, nreq=2, rate_per_s=10.0, n=(5, 5), N=(4, 4),
                             steps=(2, 2))
    with pytest.raises(bx.BfamxError) as e:
        bx.serve(dev, reqs, bx.FP32_RHS, budget_bytes=1000)
    assert e.value.code == bx.EBUDGET


def alone_state(dev, req, prec):
    sub, _ = bx.request_build(dev, req, prec)
    sub.step(req.nsteps)
    q = sub.download()
    sub.close()
    return q


@pytest.mark.parametrize("prec", [bx.FP32_RHS, bx.FP32])
def test_served_states_equal_alone_with_an_8cube_request(dev, prec):
    """The final state of every served request, handed over by the
    completion callback, equals the same request run alone bit for bit, in a
    mix that holds an 8^3 request beside smaller ones of other orders."""
    print(f"serve: final states alone vs mix with an 8^3 request, prec={prec}", flush=True)
    reqs = bx.serve_generate(seed=5, nreq=8, rate_per_s=500.0, n=(3, 5), N=(3, 7),
                             steps=(8, 20))
    reqs[2].n = 8
    states = {}
    res, st = bx.serve(dev, reqs, prec, budget_bytes=1 << 30, max_chunk=4,
                       on_done=lambda i, sub: states.__setitem__(i, sub.download()))
    assert st["max_active"] >= 3 and sorted(states) == list(range(len(reqs)))
    for i, q in enumerate(reqs):
        assert np.array_equal(states[i], alone_state(dev, q, prec)), i


@pytest.mark.parametrize("prec", [bx.FP32_RHS, bx.FP32])
def test_served_request_matches_the_oracle(dev, prec):
    """A served request's final state against the oracle's run of the same
    problem (the request's mesh, layered material, Ricker source and dt, as
    bfamx_request_inputs reports them): FP32 arithmetic to 2e-5 of the
    largest value."""
    print(f"serve: served request vs oracle, prec={prec}", flush=True)
    reqs = bx.serve_generate(seed=11, nreq=3, rate_per_s=500.0, n=(3, 4), N=(3, 5),
                             steps=(15, 25))
    states = {}
    bx.serve(dev, reqs, prec, budget_bytes=1 << 30, max_chunk=5,
             on_done=lambda i, sub: states.__setitem__(i, sub.download()))
    for i, rq in enumerate(reqs):
        inp = bx.request_inputs(rq)
        mesh = W.Mesh(3, [rq.n] * 3, rq.N)
        rho, cp, cs = inp["rho"], inp["cp"], inp["cs"]
        lam, mu = rho * (cp * cp - 2 * cs * cs), rho * cs * cs
        src = {"elem": inp["src_elem"], "weights": inp["src_weights"], "amp": inp["amp"],
               "f0": inp["f0"], "t0": inp["t0"]}
        P = oracle.OracleProblem(oracle.ELASTIC3D, rq.N, mesh.vmapM, mesh.vmapP, mesh.h, rho,
                                 lam, mu, x0=mesh.x0, src=src,
                                 ops=(mesh.r, mesh.w, 1 / mesh.w, mesh.Dr))
        want = P.run(np.zeros((9, mesh.K * mesh.Np)), inp["dt"], rq.nsteps)
        scale = np.abs(want).max()
        assert scale > 0
        assert np.abs(states[i] - want).max() < 2e-5 * scale, i

