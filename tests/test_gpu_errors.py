"""Error paths of the C ABI on a device: invalid uploads, calls out of order
and the HBM budget.  Each failing call must return its status code, leave
no device memory behind and leave the device usable for a valid upload.
The reference aborts on the first two kinds (bfam.h:36-57) and on a budget
overrun (bfamo_device_malloc, bfam.c:8561-8589); here each is a return
code."""
import numpy as np
import pytest

import bfam_amd as bx
import bfam_amd.workloads as W
from util import random_acoustic, random_elastic

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = bx.Device(0)
    yield d
    d.close()


def upload_code(dev, cfg, **override):
    m = cfg["mesh"]
    args = dict(r=m.r, w=m.w, Dr=m.Dr, vmapM=m.vmapM, vmapP=m.vmapP, h=m.h,
                rho=cfg["rho"], cp=cfg["cp"], cs=cfg.get("cs"), N=m.N)
    args.update(override)
    N = args.pop("N")
    with pytest.raises(bx.BfamxError) as e:
        bx.Subdomain(dev, cfg["physics"], bx.FP32, N, args["r"], args["w"], args["Dr"],
                     args["vmapM"], args["vmapP"], args["h"], args["rho"], args["cp"],
                     args["cs"], x0=m.x0)
    return e.value.code


def test_invalid_uploads_leave_nothing_behind(dev):
    cfg, q0 = random_elastic((2, 2, 2), 4, seed=3, src=False)
    used0, _ = dev.memory()
    K = cfg["mesh"].K
    cs0 = np.array(cfg["cs"])
    cs0[K // 2] = cfg["cp"][K // 2]  # no positive bulk modulus: cs^2 > 3/4 cp^2
    h_bad = np.array(cfg["mesh"].h)
    h_bad[1, 0] = -h_bad[1, 0]
    vP = np.array(cfg["mesh"].vmapP)
    vP[5] = K * cfg["mesh"].Np  # one past the last node
    vM = np.array(cfg["mesh"].vmapM)
    vM[[0, 1]] = vM[[1, 0]]  # not the face mask
    assert upload_code(dev, cfg, cs=cs0) == bx.EINVAL
    assert upload_code(dev, cfg, h=h_bad) == bx.EINVAL
    assert upload_code(dev, cfg, vmapP=vP) == bx.EINVAL
    assert upload_code(dev, cfg, vmapM=vM) == bx.EINVAL
    assert dev.memory()[0] == used0
    # a valid upload on the same device still steps
    sub = W.upload(dev, cfg, bx.FP32)
    sub.set_fields(q0, 0.0, cfg["dt"])
    sub.step(2)
    assert np.all(np.isfinite(sub.download()))
    sub.close()
    assert dev.memory()[0] == used0


@pytest.mark.parametrize("N", [1, 2, 8])
def test_elastic_orders_outside_3_to_7_are_unsupported(dev, N):
    mesh = W.Mesh(3, [2, 2, 2], N)
    K = mesh.K
    cfg = {"mesh": mesh, "rho": np.ones(K), "cp": np.full(K, 2.0), "cs": np.ones(K),
           "physics": bx.ELASTIC3D}
    used0, _ = dev.memory()
    assert upload_code(dev, cfg) == bx.EUNSUPPORTED
    assert dev.memory()[0] == used0


def test_calls_before_set_fields(dev):
    cfg, q0 = random_elastic((2, 2, 2), 3, seed=4, src=False)
    sub = W.upload(dev, cfg, bx.FP64)
    for call in (sub.step, sub.stage, sub.download, sub.rhs):
        with pytest.raises(bx.BfamxError) as e:
            call()
        assert e.value.code == bx.ESTATE
    sub.set_fields(q0, 0.0, cfg["dt"])
    sub.step(1)
    sub.close()


def test_budget_overrun_is_an_error_not_an_abort():
    cfg, q0 = random_elastic((4, 4, 4), 7, seed=5, src=False)
    need = bx.subdomain_bytes(3, 7, cfg["mesh"].K, bx.ELASTIC3D, bx.FP32)
    small = bx.Device(0, budget_bytes=need // 2)
    try:
        used0, budget = small.memory()
        assert budget == need // 2
        with pytest.raises(bx.BfamxError) as e:
            W.upload(small, cfg, bx.FP32)
        assert e.value.code == bx.EBUDGET
        assert small.memory()[0] == used0
    finally:
        small.close()
    big = bx.Device(0, budget_bytes=2 * need + (64 << 20))
    try:
        sub = W.upload(big, cfg, bx.FP32)
        sub.set_fields(q0, 0.0, cfg["dt"])
        sub.step(1)
        assert np.all(np.isfinite(sub.download()))
        sub.close()
    finally:
        big.close()


@pytest.mark.parametrize("prec", [bx.FP64, bx.FP64_PLAIN, bx.FP32_RHS, bx.FP32])
def test_resident_bytes_match_the_admission_estimate(dev, prec):
    """bfamx_subdomain_bytes (admission control) is what the upload
    allocates: FP32 and FP64 hold two states and no dq."""
    for physics, dims, N in ((bx.ELASTIC3D, (3, 2, 2), 3), (bx.ELASTIC3D, (2, 2, 2), 7),
                             (bx.ACOUSTIC2D, (4, 3), 5)):
        if physics == bx.ACOUSTIC2D:
            cfg, _ = random_acoustic(dims, N)
        else:
            cfg, _ = random_elastic(dims, N, seed=6, src=False)
        cfg.pop("src", None)
        used0, _ = dev.memory()
        sub = W.upload(dev, cfg, prec)
        nf, nbytes = sub.info()
        m = cfg["mesh"]
        assert nbytes == bx.subdomain_bytes(m.dim, N, m.K, physics, prec), (physics, N)
        assert dev.memory()[0] - used0 == nbytes
        sub.close()


def test_batch_members_must_share_the_stage_count(dev):
    """A batch step runs one stage count for every member: a Heun (2-stage)
    member beside KC54 (5-stage) ones is refused at batch creation."""
    cfg, q0 = random_elastic((2, 2, 2), 4, seed=9, src=False)
    kc = W.upload(dev, cfg, bx.FP32)
    heun = W.upload(dev, cfg, bx.FP32, lsrk_method=bx.LSRK_HEUN)
    with pytest.raises(bx.BfamxError) as e:
        bx.Batch(dev, [kc, heun])
    assert e.value.code == bx.EINVAL
    # two Heun members batch and step as Heun steps (2 stages each)
    heun2 = W.upload(dev, cfg, bx.FP32, lsrk_method=bx.LSRK_HEUN)
    for s in (heun, heun2):
        s.set_fields(q0, 0.0, cfg["dt"])
    b = bx.Batch(dev, [heun, heun2])
    b.step(3)
    assert heun.time()[1] == 6 and heun2.time()[1] == 6
    b.close()
    for s in (kc, heun, heun2):
        s.close()


def custom_table(a0):
    """a 2-stage table with A[0] = a0 (a nonzero A[0] needs a stored dq)"""
    t = bx.Lsrk()
    t.n_stages = 2
    t.A[0], t.A[1] = a0, -0.5
    t.B[0], t.B[1] = 1.0, 0.5
    t.C[0], t.C[1], t.C[2] = 0.0, 1.0, 1.0
    return t


@pytest.mark.parametrize("prec", [bx.FP64, bx.FP32_RHS, bx.FP32])
def test_resident_bytes_match_the_estimate_for_any_table(dev, prec):
    """bfamx_subdomain_bytes_lsrk counts the dq that a table with A[0] != 0
    makes FP32 and FP64 store, as the upload allocates it."""
    cfg, _ = random_elastic((2, 2, 2), 5, seed=10, src=False)
    m = cfg["mesh"]
    for a0 in (0.0, 0.25):
        t = custom_table(a0)
        used0, _ = dev.memory()
        sub = W.upload(dev, cfg, prec, lsrk_method=t)
        nbytes = sub.info()[1]
        assert nbytes == bx.subdomain_bytes(3, 5, m.K, bx.ELASTIC3D, prec, lsrk=t), (prec, a0)
        assert dev.memory()[0] - used0 == nbytes
        sub.close()


@pytest.mark.parametrize("case", ["permuted", "other_axis"])
def test_rotated_connectivity_is_unsupported(dev, case):
    """Two elements whose shared face is rotated (a permuted node order, or
    an x face meeting a y face, as p4est connectivities with rotated trees
    give): the kernels take axis-aligned geometry, so the upload refuses it
    instead of computing a wrong derivative.  The same pair with the
    identity orientation uploads."""
    N, dim = 3, 3
    EToE = np.tile(np.arange(2, dtype=np.int32)[:, None], (1, 6))
    EToF = np.tile(np.arange(6, dtype=np.int8)[None, :], (2, 1))
    EToE[0, 1], EToE[1, 0] = 1, 0
    if case == "permuted":
        EToF[0, 1], EToF[1, 0] = 0 + 6 * 1, 1 + 6 * 1
    else:
        EToE[1, 0], EToF[1, 0] = 1, 0  # element 1's -x face is a boundary
        EToE[1, 2] = 0
        EToF[0, 1], EToF[1, 2] = 2, 1
    vM, vP = bx.build_vmaps(dim, N, EToE, EToF)
    mesh = W.Mesh(3, [2, 1, 1], N, periodic=[False, False, False])
    K = 2
    cfg = {"mesh": mesh, "rho": np.ones(K), "cp": np.full(K, 2.0), "cs": np.ones(K),
           "physics": bx.ELASTIC3D}
    used0, _ = dev.memory()
    assert upload_code(dev, cfg, vmapM=vM, vmapP=vP) == bx.EUNSUPPORTED
    assert dev.memory()[0] == used0
    sub = W.upload(dev, cfg, bx.FP32)  # the brick's own (aligned) maps
    sub.close()
