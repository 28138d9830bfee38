"""bfamx_rhs_noalloc against bfamx_rhs through the C ABI: the same kernels
staged in the library's fixed device array instead of an allocated buffer,
so the results must agree bit for bit, on two calls in a row (the second on
a smaller subdomain, so that stale contents of the shared array would
show), with no device memory allocated by the call, and sizes beyond the
array refused."""
import numpy as np
import pytest

import bfam_amd as bx
import bfam_amd.workloads as W
from util import random_acoustic, random_elastic

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [bx.FP64, bx.FP64_PLAIN, bx.FP32_RHS, bx.FP32])
def test_rhs_noalloc_matches_rhs(prec):
    dev = bx.Device(0)
    try:
        for dims, N, seed in (((3, 3, 2), 5, 21), ((2, 2, 2), 3, 22)):
            print(f"rhs_noalloc prec={prec} dims={dims} N={N}", flush=True)
            cfg, q0 = random_elastic(dims, N, seed=seed)
            sub = W.upload(dev, cfg, prec)
            sub.set_fields(q0, 0.0, cfg["dt"])
            t = 0.3 * cfg["src"]["t0"]
            used0 = dev.memory()[0]
            got = sub.rhs_noalloc(t)
            assert dev.memory()[0] == used0  # nothing allocated
            want = sub.rhs(t)  # allocates its staging buffer on first use
            assert np.array_equal(got, want)
            sub.step(1)
            assert np.array_equal(sub.rhs_noalloc(2 * t), sub.rhs(2 * t))
            sub.close()
        cfg, q0 = random_acoustic((5, 4), 6)
        sub = W.upload(dev, cfg, prec)
        sub.set_fields(q0, 0.0, cfg["dt"])
        assert np.array_equal(sub.rhs_noalloc(0.1), sub.rhs(0.1))
        sub.close()
    finally:
        dev.close()


def test_rhs_noalloc_refuses_sizes_beyond_the_scratch():
    dev = bx.Device(0)
    try:
        # 18 x 18 x 16 elements at N = 7: 9 * 5184 * 512 = 23.9 M values > 8.39 M
        cfg, q0 = random_elastic((18, 18, 16), 7, seed=23, src=False)
        sub = W.upload(dev, cfg, bx.FP32)
        sub.set_fields(q0, 0.0, cfg["dt"])
        with pytest.raises(bx.BfamxError) as e:
            sub.rhs_noalloc(0.0)
        assert e.value.code == bx.EINVAL
        sub.close()
    finally:
        dev.close()
