"""Serving through the C ABI: requests admitted as they arrive under an HBM
budget and advanced together by one batched launch per stage.  Each
request's result is the same alone and in the mix, bit for bit (every
element is computed by the same code whatever else shares the launch)."""
import numpy as np
import pytest

import bfam_amd as bx

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
