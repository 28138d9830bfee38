"""Adaptation on the device: bfamx_transfer_fields moves the state of one
adapted mesh's subdomains to another's (h-refine, h-coarsen and the order
change between levels) and is compared with the oracle's transfer of the
same source state: FP64 to 1e-13, FP32 state to 1e-6 (max norm relative to
the largest value; the arithmetic is FP64, each value rounded once to its
state type).  The sources are taken after one LSRK step, so that they are
read from the second ping-pong buffer; the glued destination mesh then
steps on from the transferred state as the oracle does."""
import numpy as np
import pytest

import bfam_amd as bx
import bfam_amd.workloads as W
import oracle
from util import oracle_multi

pytestmark = pytest.mark.gpu

TOL = {bx.FP64: 1e-13, bx.FP32_RHS: 1e-13, bx.FP32: 1e-6}


@pytest.fixture(scope="module")
def dev():
    d = bx.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def pair():
    return W.adaptation_pair()


def rel_max(got, want):
    return max(np.abs(g - w).max() for g, w in zip(got, want)) / \
        max(np.abs(w).max() for w in want)


def oracle_transfer(a, b, m, qs):
    return oracle.transfer(m, qs, [sm.N for sm in a["subs"]], [sm.K for sm in b["subs"]],
                           [sm.N for sm in b["subs"]])


@pytest.mark.parametrize("prec", [bx.FP64, bx.FP32_RHS, bx.FP32])
def test_transfer_matches_the_oracle(dev, pair, prec):
    print(f"transfer vs oracle prec={prec}", flush=True)
    a, b, m = pair
    rng = np.random.default_rng(11)
    q0 = [rng.standard_normal((9, sm.K * sm.Np)) for sm in a["subs"]]
    src = W.upload_multi(dev, a, prec)
    for s, q in zip(src, q0):
        s.set_fields(q, 0.0, a["dt"])
    glue = bx.Glue(dev, src, a["mortars"])
    batch = bx.Batch(dev, src)
    batch.set_glue(glue)
    batch.step(1)  # the state now sits in the second buffer
    batch.close()
    glue.close()
    t_src = src[0].time()[0]
    before = [s.download() for s in src]
    dst = W.upload_multi(dev, b, prec)
    bx.transfer_fields(m, src, dst, 0.5 * b["dt"])
    got = [s.download() for s in dst]
    times = [s.time() for s in dst]
    for s in src + dst:
        s.close()
    want = oracle_transfer(a, b, m, before)
    assert rel_max(got, want) < TOL[prec]
    assert all(t == t_src and st == 0 for t, st in times)


def test_transfer_moves_polynomials_exactly_on_the_device(dev, pair):
    print("transfer of a polynomial field", flush=True)
    from test_transfer import poly_fields
    a, b, m = pair
    src = W.upload_multi(dev, a, bx.FP64)
    for s, q in zip(src, poly_fields(a["subs"], 3)):
        s.set_fields(q, 0.0, a["dt"])
    dst = W.upload_multi(dev, b, bx.FP64)
    bx.transfer_fields(m, src, dst, b["dt"])
    got = [s.download() for s in dst]
    for s in src + dst:
        s.close()
    want = poly_fields(b["subs"], 3)
    assert rel_max(got, want) < 1e-12


def test_adapted_mesh_steps_on_after_the_transfer(dev, pair):
    """Transfer, glue the new mesh, take two steps: the oracle, given the
    oracle's transfer of the same state, ends at the same state."""
    print("step after transfer", flush=True)
    a, b, m = pair
    rng = np.random.default_rng(12)
    q0 = [rng.standard_normal((9, sm.K * sm.Np)) for sm in a["subs"]]
    src = W.upload_multi(dev, a, bx.FP64)
    for s, q in zip(src, q0):
        s.set_fields(q, 0.0, a["dt"])
    dst = W.upload_multi(dev, b, bx.FP64)
    bx.transfer_fields(m, src, dst, b["dt"])
    glue = bx.Glue(dev, dst, b["mortars"])
    batch = bx.Batch(dev, dst)
    batch.set_glue(glue)
    batch.step(2)
    got = [s.download() for s in dst]
    batch.close()
    glue.close()
    for s in src + dst:
        s.close()
    mid = oracle_transfer(a, b, m, q0)
    want = oracle_multi(b).run(mid, b["dt"], 2)
    assert rel_max(got, want) < 1e-12


def test_transfer_error_paths(dev, pair):
    """A source inside a step, sources at different times, maps that name a
    destination element twice: status codes, nothing written."""
    print("transfer error paths", flush=True)
    a, b, m = pair
    src = W.upload_multi(dev, a, bx.FP64)
    for s, sm in zip(src, a["subs"]):
        s.set_fields(np.zeros((9, sm.K * sm.Np)), 0.0, a["dt"])
    dst = W.upload_multi(dev, b, bx.FP64)
    with pytest.raises(bx.BfamxError) as e:
        bx.transfer_fields(m, src[:1], dst, b["dt"])  # names sources that are not given
    assert e.value.code == bx.EINVAL
    bad = dict(m)
    bad["dst_elem"] = m["dst_elem"].copy()
    bad["dst_elem"][1] = bad["dst_elem"][0]
    bad["dst_subd"] = m["dst_subd"].copy()
    bad["dst_subd"][1] = bad["dst_subd"][0]
    with pytest.raises(bx.BfamxError) as e:
        bx.transfer_fields(bad, src, dst, b["dt"])
    assert e.value.code == bx.EINVAL
    src[0].set_fields(np.zeros((9, a["subs"][0].K * a["subs"][0].Np)), 1.0, a["dt"])
    with pytest.raises(bx.BfamxError) as e:
        bx.transfer_fields(m, src, dst, b["dt"])
    assert e.value.code == bx.ESTATE
    src[0].set_fields(np.zeros((9, a["subs"][0].K * a["subs"][0].Np)), 0.0, a["dt"])
    src[1].stage()  # inside a step
    with pytest.raises(bx.BfamxError) as e:
        bx.transfer_fields(m, src, dst, b["dt"])
    assert e.value.code == bx.ESTATE
    for s in src + dst:
        s.close()
