"""Device parity against the CPU FP64 oracle (oracle/oracle.c), through the
C ABI.  Tolerances: FP64 paths 1e-12 relative (max norm over all fields),
FP32 arithmetic 2e-5 relative; results that must not depend on launch shape
(batching, graphs, XCD remapping) are compared bit for bit."""
import numpy as np
import pytest

import bfam_amd as bx
import bfam_amd.workloads as W
from util import oracle_problem, random_acoustic, random_elastic, rel_l2, rel_max

pytestmark = pytest.mark.gpu

TOL = {bx.FP64: 1e-12, bx.FP64_PLAIN: 1e-12, bx.FP32_RHS: 2e-5, bx.FP32: 2e-5}
PRECS = [bx.FP64, bx.FP64_PLAIN, bx.FP32_RHS, bx.FP32]


@pytest.fixture(scope="module")
def dev():
    d = bx.Device(0)
    yield d
    d.close()


@pytest.fixture(params=[0, 1, 2, 3, 4],
                ids=["line", "node", "node_resident", "node_alt", "node_alt_resident"])
def kernel(request, dev):
    """run with the line (0) or the node-parallel elastic stage kernel: one
    block per task (1) or on a resident grid (2), and the same with its
    other elements-per-block at N = 3..5 (3, 4)"""
    for prec in (bx.FP64, bx.FP32_RHS, bx.FP32):
        for N in range(3, 8):
            dev.set_kernel(prec, N, request.param)
    yield request.param
    for prec in (bx.FP64, bx.FP32_RHS, bx.FP32):
        for N in range(3, 8):
            dev.set_kernel(prec, N, -1)


@pytest.mark.parametrize("N", [3, 4, 5, 6, 7])
@pytest.mark.parametrize("prec", PRECS)
def test_elastic_rhs(dev, kernel, N, prec):
    print(f"elastic rhs N={N} prec={prec} kernel={kernel}", flush=True)
    cfg, q0 = random_elastic((3, 2, 4), N, seed=N)
    sub = W.upload(dev, cfg, prec)
    sub.set_fields(q0, 0.0, cfg["dt"])
    got = sub.rhs(0.25)
    sub.close()
    want = oracle_problem(cfg).rhs(q0, 0.25)
    assert rel_max(got, want) < TOL[prec]


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("prec", PRECS)
def test_acoustic_rhs(dev, N, prec):
    print(f"acoustic rhs N={N} prec={prec}", flush=True)
    cfg, q0 = random_acoustic((5, 3), N, seed=N)
    sub = W.upload(dev, cfg, prec)
    sub.set_fields(q0, 0.0, cfg["dt"])
    got = sub.rhs(0.05)
    sub.close()
    want = oracle_problem(cfg).rhs(q0, 0.05)
    assert rel_max(got, want) < TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
def test_elastic_lsrk_steps(dev, kernel, prec):
    print(f"elastic lsrk 4 steps prec={prec} kernel={kernel}", flush=True)
    cfg, q0 = random_elastic((4, 4, 3), 5, seed=3)
    sub = W.upload(dev, cfg, prec)
    sub.set_fields(q0, 0.0, cfg["dt"])
    sub.step(4)
    t, st = sub.time()
    got = sub.download()
    sub.close()
    want = oracle_problem(cfg).run(q0, cfg["dt"], 4)
    assert st == 20 and abs(t - 4 * cfg["dt"]) < 1e-15
    assert rel_max(got, want) < TOL[prec]


@pytest.mark.parametrize("method", ["W33", "HEUN", "FE"])
@pytest.mark.parametrize("prec", [bx.FP32, bx.FP64, bx.FP32_RHS])
def test_other_lsrk_tables(dev, method, prec):
    """The other low-storage tables step as the oracle does; FP32 and FP64
    rebuild dq from the states with each table's own A_s / B_{s-1}."""
    import oracle
    print(f"elastic lsrk {method} prec={prec}", flush=True)
    cfg, q0 = random_elastic((4, 4, 3), 5, seed=9)
    m = cfg["mesh"]
    sub = bx.Subdomain(dev, bx.ELASTIC3D, prec, m.N, m.r, m.w, m.Dr, m.vmapM, m.vmapP, m.h,
                       cfg["rho"], cfg["cp"], cfg["cs"], x0=m.x0,
                       lsrk_method=getattr(bx, "LSRK_" + method))
    s = cfg["src"]
    sub.set_source(s["elem"], s["weights"], s["amp"], s["f0"], s["t0"])
    sub.set_fields(q0, 0.0, cfg["dt"])
    sub.step(3)
    t, st = sub.time()
    got = sub.download()
    sub.close()
    want = oracle_problem(cfg).run(q0, cfg["dt"], 3, method=getattr(oracle, "LSRK_" + method))
    ns = {"W33": 3, "HEUN": 2, "FE": 1}[method]
    assert st == 3 * ns and abs(t - 3 * cfg["dt"]) < 1e-15
    assert rel_max(got, want) < TOL[prec]


def test_acoustic_plane_wave(dev):
    """Config acoustic2d_plane_wave: 100 KC54 steps, FP64, against the oracle
    and against the exact plane wave."""
    print("acoustic plane wave 16x16 N=4", flush=True)
    cfg = W.acoustic2d_plane_wave()
    sub = W.upload(dev, cfg, bx.FP64)
    sub.set_fields(cfg["q0"], 0.0, cfg["dt"])
    sub.step(cfg["nsteps"])
    got = sub.download()
    E = sub.energy()
    sub.close()
    P = oracle_problem(cfg)
    want = P.run(cfg["q0"], cfg["dt"], cfg["nsteps"])
    assert rel_max(got, want) < 1e-12
    ex = W.acoustic2d_exact(cfg, cfg["dt"] * cfg["nsteps"])
    assert rel_l2(got, ex) < 2e-6
    assert abs(E - P.energy(want)) < 1e-12 * P.energy(want)


def test_energy_and_l2(dev):
    cfg, q0 = random_elastic((3, 3, 3), 4, seed=9)
    a = W.upload(dev, cfg, bx.FP64)
    b = W.upload(dev, cfg, bx.FP32)
    a.set_fields(q0)
    b.set_fields(q0)
    P = oracle_problem(cfg)
    assert abs(a.energy() - P.energy(q0)) < 1e-12 * P.energy(q0)
    d2, n2 = b.l2_diff(a)
    assert n2 > 0 and np.sqrt(d2 / n2) < 1e-6
    a.close()
    b.close()


def test_batch_equals_single(dev):
    """Same-order batch and mixed-order batch reproduce single launches bit
    for bit (each element is computed by the same code)."""
    print("batch vs single", flush=True)
    subs, ref = [], []
    for i, N in enumerate([3, 5, 7, 5]):
        cfg, q0 = random_elastic((2, 3, 2), N, seed=20 + i)
        s = W.upload(dev, cfg, bx.FP32_RHS)
        s.set_fields(q0, 0.0, cfg["dt"])
        s.step(3)
        ref.append(s.download())
        s.set_fields(q0, 0.0, cfg["dt"])
        subs.append(s)
    b = bx.Batch(dev, subs)
    b.step(3)
    for s, r in zip(subs, ref):
        assert np.array_equal(s.download(), r)
    b.close()
    for s in subs:
        s.close()


def test_graph_replay_matches_eager(dev):
    print("graph replay", flush=True)
    cfg, q0 = random_elastic((4, 2, 3), 6, seed=31)
    a = W.upload(dev, cfg, bx.FP64)
    a.set_fields(q0, 0.0, cfg["dt"])
    a.step(3)
    eager = a.download()
    a.set_fields(q0, 0.0, cfg["dt"])
    b = bx.Batch(dev, [a])
    b.capture(1)
    outs = []
    for _ in range(3):
        b.launch_graph()
        dev.sync()
        outs.append(a.download())
    assert np.array_equal(outs[-1], eager)
    assert not np.array_equal(outs[0], outs[1])
    b.close()
    a.close()


@pytest.mark.parametrize("N", [3, 7])
def test_occupancy_caps_are_bit_identical(dev, kernel, N):
    """Capping the resident blocks per CU (dynamic LDS the kernel does not
    use) changes residency only: every level gives the same bits, and the
    autotuner leaves a valid kernel and level for the order."""
    print(f"occupancy caps N={N} kernel={kernel}", flush=True)
    cfg, q0 = random_elastic((4, 3, 3), N, seed=50 + N)
    res = []
    # the resident grid shares the cap of its kernel and block size
    kind = {0: 0, 1: 1, 2: 1, 3: 2, 4: 2}[kernel]
    for lv in (0, 1, 2):
        dev.set_occupancy(bx.FP32, N, kind, lv)
        assert dev.get_occupancy(bx.FP32, N, kind) == lv
        s = W.upload(dev, cfg, bx.FP32)
        s.set_fields(q0, 0.0, cfg["dt"])
        s.step(2)
        res.append(s.download())
        s.close()
    dev.set_occupancy(bx.FP32, N, kind, 0)
    assert np.array_equal(res[0], res[1]) and np.array_equal(res[0], res[2])
    with pytest.raises(bx.BfamxError):
        dev.set_occupancy(bx.FP32, N, kind, 3)


@pytest.mark.parametrize("N,n", [(3, 18), (4, 14), (5, 12), (7, 12)])
@pytest.mark.parametrize("prec", [bx.FP32, bx.FP64])
def test_resident_grid_is_bit_identical(dev, N, n, prec):
    """The node-parallel kernel on a resident grid (each block looping over
    tasks; these bricks have more tasks than blocks fit the CUs at once)
    gives the bits of one block per task, and so does either with the
    other elements-per-block, except FP64 at N = 4, whose two block sizes
    agree to the last bits"""
    print(f"resident grid N={N} n={n} prec={prec}", flush=True)
    cfg = W.elastic3d_ricker(n=n, N=N, nsteps=20)
    m = cfg["mesh"]
    q0 = np.random.default_rng(N).standard_normal((9, m.K * m.Np))
    res = []
    for k in (1, 2, 3, 4):
        dev.set_kernel(prec, N, k)
        assert dev.get_kernel(prec, N) == k
        s = W.upload(dev, cfg, prec)
        s.set_fields(q0, 0.0, cfg["dt"])
        s.step(2)
        res.append(s.download())
        s.close()
    dev.set_kernel(prec, N, -1)
    if prec == bx.FP64 and N == 4:
        # the four-element FP64 blocks of N = 4 (kinds 3, 4) differ from the
        # one-element ones in the last bits only
        assert np.array_equal(res[0], res[1])
        d = [rel_max(r, res[0]) for r in res[2:]]
        print(f"FP64 N=4 four-element blocks: rel max {d}", flush=True)
        assert max(d) < 1e-13
        return
    for r in res[1:]:
        assert np.array_equal(res[0], r)


@pytest.mark.parametrize("N", [3, 4])
def test_element_blocks_bit_identical_in_every_block_order(dev, N):
    """FP32 at N = 3, 4 launches one block per element of each node task
    (block v takes element v % T of task v / T) where a block per task takes
    the task's elements in turn: every block order, and the resident grid
    that takes whole tasks, give the same bits.  The 5x5x3 brick's last
    task is partial, so some blocks find no element."""
    print(f"element blocks N={N}", flush=True)
    cfg, q0 = random_elastic((5, 5, 3), N, seed=60 + N)
    res = []
    for remap, kind in ((2, 1), (0, 1), (1, 1), (2, 2)):
        dev.set_tuning(remap)
        dev.set_kernel(bx.FP32, N, kind)
        s = W.upload(dev, cfg, bx.FP32)
        s.set_fields(q0, 0.0, cfg["dt"])
        s.step(2)
        res.append(s.download())
        s.close()
    dev.set_tuning(2)
    dev.set_kernel(bx.FP32, N, -1)
    for r in res[1:]:
        assert np.array_equal(res[0], r)


def test_xcd_remap_invariant(dev):
    """Every block order gives the same bits (at N=7 the 5x5x3 brick has one
    whole group of 64 elements and a partial one in mode 2, the 8x8x2 brick
    two whole groups)."""
    cfg, q0 = random_elastic((5, 5, 3), 7, seed=41)
    res = []
    for remap in (1, 0, 2):
        dev.set_tuning(remap)
        s = W.upload(dev, cfg, bx.FP32)
        s.set_fields(q0, 0.0, cfg["dt"])
        s.step(2)
        res.append(s.download())
        s.close()
    cfg, q0 = random_elastic((8, 8, 2), 7, seed=42)
    res2 = []
    for remap in (0, 2):
        dev.set_tuning(remap)
        s = W.upload(dev, cfg, bx.FP32)
        s.set_fields(q0, 0.0, cfg["dt"])
        s.step(2)
        res2.append(s.download())
        s.close()
    dev.set_tuning(2)
    assert np.array_equal(res[0], res[1]) and np.array_equal(res[0], res[2])
    assert np.array_equal(res2[0], res2[1])


def test_mms_forcing(dev):
    """Manufactured solution: device RHS with forcing equals the oracle's, and
    a few steps track the exact solution."""
    print("mms 4^3 N=5", flush=True)
    cfg = W.elastic3d_mms(n=4, N=5)
    m = cfg["mesh"]
    q0 = W.mms_exact(m, cfg["A"], cfg["k"], cfg["omega"], 0.0)
    sub = W.upload(dev, cfg, bx.FP64)
    sub.set_fields(q0, 0.0, cfg["dt"])
    got = sub.rhs(0.1)
    P = oracle_problem(cfg)
    assert rel_max(got, P.rhs(q0, 0.1)) < 1e-12
    sub.step(10)
    d2, n2 = sub.mms_error(cfg["A"])
    q = sub.download()
    sub.close()
    want = P.run(q0, cfg["dt"], 10)
    assert rel_max(q, want) < 1e-12
    assert np.sqrt(d2 / n2) < 1e-3


@pytest.mark.parametrize("prec", [bx.FP64, bx.FP32_RHS, bx.FP32])
def test_mms_order_7(dev, kernel, prec):
    """The precision-study configuration's order (N = 7) on a 3^3 brick: the
    forced RHS and 10 LSRK4(5) steps of the manufactured solution against
    the oracle in every fused precision and with both stage kernels (the
    one-element-per-block kernel instances), and the error against the exact
    solution."""
    print(f"mms 3^3 N=7 prec={prec} kernel={kernel}", flush=True)
    cfg = W.elastic3d_mms(n=3, N=7)
    m = cfg["mesh"]
    q0 = W.mms_exact(m, cfg["A"], cfg["k"], cfg["omega"], 0.0)
    P = oracle_problem(cfg)
    sub = W.upload(dev, cfg, prec)
    sub.set_fields(q0, 0.0, cfg["dt"])
    assert rel_max(sub.rhs(0.1), P.rhs(q0, 0.1)) < TOL[prec]
    sub.step(10)
    d2, n2 = sub.mms_error(cfg["A"])
    q = sub.download()
    sub.close()
    assert rel_max(q, P.run(q0, cfg["dt"], 10)) < TOL[prec]
    assert np.sqrt(d2 / n2) < 1e-3


@pytest.mark.parametrize("N", [3, 5, 7])
def test_bf16x3_matrix_core_derivatives(dev, N):
    """Split-BF16x3 derivatives on the matrix cores: the RHS and a few steps
    against the oracle within the accuracy two-term bf16 operands give
    (2e-4 of the largest value), and visibly less accurate than FP32 (so
    that the matrix-core path is the one that ran)."""
    print(f"bf16x3 N={N}", flush=True)
    cfg, q0 = random_elastic((3, 2, 2), N, seed=40 + N)
    P = oracle_problem(cfg)
    t = 0.2 * cfg["src"]["t0"]
    want = P.rhs(q0, t)
    errs = {}
    for prec in (bx.FP32, bx.FP32_BF16X3):
        sub = W.upload(dev, cfg, prec)
        sub.set_fields(q0, 0.0, cfg["dt"])
        errs[prec] = rel_max(sub.rhs(t), want)
        sub.step(3)
        q = sub.download()
        sub.close()
        assert rel_max(q, P.run(q0, cfg["dt"], 3)) < 2e-4
    assert errs[bx.FP32_BF16X3] < 2e-4
    assert errs[bx.FP32_BF16X3] > 4 * errs[bx.FP32]


def test_bf16x3_manufactured_solution_order_7(dev):
    print("bf16x3 mms 3^3 N=7", flush=True)
    cfg = W.elastic3d_mms(n=3, N=7)
    m = cfg["mesh"]
    q0 = W.mms_exact(m, cfg["A"], cfg["k"], cfg["omega"], 0.0)
    P = oracle_problem(cfg)
    sub = W.upload(dev, cfg, bx.FP32_BF16X3)
    sub.set_fields(q0, 0.0, cfg["dt"])
    sub.step(10)
    d2, n2 = sub.mms_error(cfg["A"])
    q = sub.download()
    sub.close()
    assert rel_max(q, P.run(q0, cfg["dt"], 10)) < 2e-4
    assert np.sqrt(d2 / n2) < 1e-3


def test_bf16x3_runs_the_node_kernel_only(dev):
    with pytest.raises(bx.BfamxError) as e:
        dev.set_kernel(bx.FP32_BF16X3, 7, 0)
    assert e.value.code == bx.EUNSUPPORTED
    assert dev.get_kernel(bx.FP32_BF16X3, 5) == 1
    cfg, _ = random_acoustic((3, 3), 4)
    with pytest.raises(bx.BfamxError) as e:
        W.upload(dev, cfg, bx.FP32_BF16X3)
    assert e.value.code == bx.EUNSUPPORTED


def test_full_size_rhs_against_oracle(dev):
    """The 48^3 N=7 elastic brick at full size: one FP64 right-hand side
    evaluation (with the Ricker source active) against the oracle on the
    host's cores."""
    print("full size 48^3 N=7 rhs vs oracle", flush=True)
    cfg = W.elastic3d_ricker()
    m = cfg["mesh"]
    rng = np.random.default_rng(12)
    q0 = rng.standard_normal((9, m.K * m.Np))
    t = 0.5 * cfg["src"]["t0"]
    sub = W.upload(dev, cfg, bx.FP64)
    sub.set_fields(q0, 0.0, cfg["dt"])
    got = sub.rhs(t)
    sub.close()
    want = oracle_problem(cfg).rhs(q0, t)
    assert rel_max(got, want) < 1e-12


def test_full_size_fp32_accuracy(dev):
    """48^3 N=7, 200 LSRK4(5) steps from rest with the Ricker source: the
    FP32 path stays within relative L2 1e-5 of the FP64 path (which matches
    the oracle to 1e-12), and FP32 state with FP32 rates is no worse than an
    order of magnitude beyond FP64-state / FP32-rate."""
    print("full size 48^3 N=7 fp32 vs fp64, 200 steps", flush=True)
    cfg = W.elastic3d_ricker()
    m = cfg["mesh"]
    q0 = np.zeros((9, m.K * m.Np))
    subs = {p: W.upload(dev, cfg, p) for p in (bx.FP64, bx.FP32, bx.FP32_RHS)}
    for s in subs.values():
        s.set_fields(q0, 0.0, cfg["dt"])
        s.step(cfg["nsteps"])
    d2, n2 = subs[bx.FP32].l2_diff(subs[bx.FP64])
    e32 = np.sqrt(d2 / n2)
    d2, n2 = subs[bx.FP32_RHS].l2_diff(subs[bx.FP64])
    e32r = np.sqrt(d2 / n2)
    print(f"rel L2: fp32 {e32:.3e}, fp32_rhs {e32r:.3e}", flush=True)
    for s in subs.values():
        s.close()
    assert n2 > 0
    assert e32 < 1e-5 and e32r < 1e-5


@pytest.mark.parametrize("phys", ["elastic", "acoustic"])
def test_cfl_dt_reduction(dev, phys):
    if phys == "elastic":
        cfg, _ = random_elastic((3, 2, 4), 5, seed=2)
    else:
        cfg, _ = random_acoustic((4, 3), 4, seed=2)
    m = cfg["mesh"]
    sub = W.upload(dev, cfg, bx.FP32)
    dt = sub.cfl_dt(0.5)
    sub.close()
    want = 0.5 * np.min(m.h.min(axis=1) / cfg["cp"]) / (m.N + 1) ** 2
    assert abs(dt - want) <= 1e-15 * want
    assert abs(dt - cfg["dt"]) <= 1e-15 * want


@pytest.mark.parametrize("prec", [bx.FP32_RHS, bx.FP64])
def test_mixed_order_launch_strategies(dev, prec):
    """Orders 3..7 in one batch: the order-switching kernel and one launch
    per order give the same bits, and autotuning (which times real launches
    on the batch) leaves the state unchanged."""
    print(f"mixed vs grouped prec={prec}", flush=True)
    # the order-switching kernel is the line kernel: compare line launches
    for N in range(3, 8):
        dev.set_kernel(prec, N, 0)
    res = []
    for strategy in (0, 1, "auto"):
        subs = []
        for i, N in enumerate([3, 4, 5, 6, 7]):
            cfg, q0 = random_elastic((2, 2, 3), N, seed=60 + i)
            s = W.upload(dev, cfg, prec)
            s.set_fields(q0, 0.0, cfg["dt"])
            subs.append(s)
        b = bx.Batch(dev, subs)
        if strategy == "auto":
            assert b.autotune(trials=2, force=True) in (0, 1)
            # compare on the line kernels: the autotuner must not have
            # touched the state
            for N in range(3, 8):
                dev.set_kernel(prec, N, 0)
        else:
            b.set_strategy(strategy)
        b.step(2)
        res.append([s.download() for s in subs])
        b.close()
        for s in subs:
            s.close()
    for N in range(3, 8):
        dev.set_kernel(prec, N, -1)
    for r in res[1:]:
        for a, c in zip(res[0], r):
            assert np.array_equal(a, c)


@pytest.mark.parametrize("prec", [bx.FP32, bx.FP64, bx.FP32_RHS])
def test_autotune_mid_step_keeps_state(dev, prec):
    """Autotuning in the middle of a step (its trial launches repeat the
    current stage) restores what they overwrite: dq, or, where dq is rebuilt
    from the states, the previous state.  The step then continues bit for
    bit as without autotuning."""
    print(f"autotune mid-step prec={prec}", flush=True)
    res = []
    for tune in (False, True):
        cfg, q0 = random_elastic((3, 2, 2), 7, seed=77)
        s = W.upload(dev, cfg, prec)
        s.set_fields(q0, 0.0, cfg["dt"])
        s.stage()
        s.stage()
        b = bx.Batch(dev, [s])
        if tune:
            b.autotune(trials=2, force=True)
        dev.set_kernel(prec, 7, 1)
        b.step(1)
        res.append(s.download())
        b.close()
        s.close()
    dev.set_kernel(prec, 7, -1)
    assert np.array_equal(res[0], res[1])


def test_mms_energy_error_and_precisions(dev):
    """The device energy-norm error against the manufactured solution equals
    the oracle's energy of (q - exact); FP32_RHS (FP64 state) and FP32 stay
    close to FP64 over 100 steps."""
    print("mms energy error + precisions 5^3 N=5", flush=True)
    cfg = W.elastic3d_mms(n=5, N=5)
    m = cfg["mesh"]
    q0 = W.mms_exact(m, cfg["A"], cfg["k"], cfg["omega"], 0.0)
    P = oracle_problem(cfg)
    out = {}
    for prec in (bx.FP64, bx.FP32_RHS, bx.FP32):
        s = W.upload(dev, cfg, prec)
        s.set_fields(q0, 0.0, cfg["dt"])
        s.step(100)
        q = s.download()
        t = s.time()[0]
        ed, en = s.mms_energy_error(cfg["A"])
        ex = W.mms_exact(m, cfg["A"], cfg["k"], cfg["omega"], t)
        assert abs(ed - P.energy(q - ex)) <= 1e-10 * P.energy(ex)
        assert abs(en - P.energy(ex)) <= 1e-12 * P.energy(ex)
        out[prec] = q
        s.close()
    ref = out[bx.FP64]
    scale = np.sqrt((ref ** 2).sum())
    assert np.sqrt(((out[bx.FP32_RHS] - ref) ** 2).sum()) / scale < 1e-5
    assert np.sqrt(((out[bx.FP32] - ref) ** 2).sum()) / scale < 1e-5


@pytest.mark.parametrize("prec", [bx.FP32, bx.FP64, bx.FP32_RHS])
def test_persistent_launch_invariant(dev, prec):
    """Persistent wavefronts (tasks looped, next state streamed into LDS
    during the current output) give the same bits as one wavefront per
    task, for an element count that leaves partial last tasks (N=4: 4
    elements per wavefront)."""
    res = []
    for N in (7, 4):
        dev.set_kernel(prec, N, 0)  # persistence is a line-kernel mode
    for on in (1, 0):
        dev.set_persistent(on)
        for N in (7, 4):
            cfg, q0 = random_elastic((5, 3, 3) if N == 7 else (5, 3, 5), N, seed=70 + N)
            s = W.upload(dev, cfg, prec)
            s.set_fields(q0, 0.0, cfg["dt"])
            s.step(2)
            res.append(s.download())
            if on:
                want = oracle_problem(cfg).run(q0, cfg["dt"], 2)
                assert rel_max(res[-1], want) < TOL[prec]
            s.close()
    dev.set_persistent(0)
    for N in (7, 4):
        dev.set_kernel(prec, N, -1)
    assert np.array_equal(res[0], res[2]) and np.array_equal(res[1], res[3])


def bc_elastic(N, seed):
    rng = np.random.default_rng(seed)
    mesh = W.Mesh(3, [3, 2, 3], N, L=[1.0, 0.7, 1.2], periodic=[0, 0, 1])
    K = mesh.K
    rho = rng.uniform(1.5, 3.0, K)
    cp = rng.uniform(3.0, 6.0, K)
    cs = cp / rng.uniform(1.6, 2.0, K)
    cfg = {"mesh": mesh, "rho": rho, "cp": cp, "cs": cs, "physics": bx.ELASTIC3D,
           "dt": W.cfl_dt(float(mesh.h.min()), float(cp.max()), N, 0.5)}
    cfg["bc"] = W.boundary_kinds(mesh, [bx.BC_RIGID, bx.BC_FREE, bx.BC_ABSORB, bx.BC_FREE, 0, 0])
    return cfg, rng.standard_normal((9, K * mesh.Np))


@pytest.mark.parametrize("N", [4, 7])
@pytest.mark.parametrize("prec", PRECS)
def test_elastic_boundaries_rhs(dev, kernel, N, prec):
    """Free-surface, rigid and absorbing faces against the oracle."""
    print(f"elastic boundaries N={N} prec={prec}", flush=True)
    cfg, q0 = bc_elastic(N, 80 + N)
    assert (cfg["bc"] > 0).sum() > 0
    s = W.upload(dev, cfg, prec)
    s.set_fields(q0, 0.0, cfg["dt"])
    got = s.rhs(0.0)
    s.close()
    want = oracle_problem(cfg).rhs(q0, 0.0)
    assert rel_max(got, want) < TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
def test_acoustic_boundaries_rhs_and_steps(dev, prec):
    print(f"acoustic boundaries prec={prec}", flush=True)
    rng = np.random.default_rng(5)
    mesh = W.Mesh(2, [4, 3], 5, L=[1.0, 0.8], periodic=[0, 0])
    K = mesh.K
    cfg = {"mesh": mesh, "rho": rng.uniform(1, 2, K), "cp": rng.uniform(1, 3, K), "cs": None,
           "physics": bx.ACOUSTIC2D}
    cfg["dt"] = W.cfl_dt(float(mesh.h.min()), float(cfg["cp"].max()), 5, 0.5)
    cfg["bc"] = W.boundary_kinds(mesh, [bx.BC_FREE, bx.BC_RIGID, bx.BC_ABSORB, bx.BC_RIGID])
    q0 = rng.standard_normal((3, K * mesh.Np))
    s = W.upload(dev, cfg, prec)
    s.set_fields(q0, 0.0, cfg["dt"])
    got = s.rhs(0.0)
    s.set_fields(q0, 0.0, cfg["dt"])
    s.step(3)
    q3 = s.download()
    s.close()
    P = oracle_problem(cfg)
    assert rel_max(got, P.rhs(q0, 0.0)) < TOL[prec]
    assert rel_max(q3, P.run(q0, cfg["dt"], 3)) < TOL[prec]


def test_boundary_rejects_interior_face(dev):
    cfg, q0 = random_elastic((2, 2, 2), 3, seed=1)
    s = W.upload(dev, cfg, bx.FP64)
    with pytest.raises(bx.BfamxError) as e:
        s.set_boundary(np.ones((cfg["mesh"].K, 6), np.int8))
    assert e.value.code == bx.EINVAL
    s.close()


def test_default_kernel_policy():
    """Until autotuned, each precision and order runs the kernel that won its
    row of the per-order measurements; set_kernel overrides it and -1 goes
    back to that default."""
    d = bx.Device(0)
    try:
        assert [d.get_kernel(bx.FP32, N) for N in range(3, 8)] == [1, 2, 1, 2, 1]
        assert all(d.get_kernel(bx.FP32_RHS, N) == 1 for N in range(3, 8))
        assert [d.get_kernel(bx.FP64, N) for N in range(3, 8)] == [1, 1, 1, 2, 2]
        assert all(d.get_kernel(bx.FP32_BF16X3, N) == 1 for N in range(3, 8))
        d.set_kernel(bx.FP32, 7, 0)
        assert d.get_kernel(bx.FP32, 7) == 0
        d.set_kernel(bx.FP32, 7, -1)
        assert d.get_kernel(bx.FP32, 7) == 1
    finally:
        d.close()
