#!/usr/bin/env python3
"""Headline benchmark: DOF-stage-updates/s of the fused FP32 elastic DG stage
against the plain FP64 GPU form, on a 48^3 periodic hex brick at LGL order 7
(9 fields, 509,607,936 DOF), PCG32 per-element material, Ricker point source,
LSRK4(5) (Carpenter-Kennedy 5-stage, 2N storage).

A step is one LSRK step (5 stages) over the whole brick.  The FP32 run goes
through the batched path (one launch per stage over all resident subdomains,
the step captured in a HIP graph and replayed).  Inputs are resident in HBM
before the timed region; the clock is a pair of HIP events on the stream the
kernels run on.

The JSON line also carries:
  fp64 / fp64_plain  rates of the fused FP64 stage and of the unfused FP64
                     form (volume+surface kernel, then the update kernel);
  accuracy           relative L2 of FP32 against FP64 after 200 steps from
                     the same initial state (LGL-quadrature weighted, all
                     nine fields);
  roofline           the FP32 stage kernel against HBM: algorithmic bytes
                     (q read, previous q read, next q written, 4 B each =
                     12 B per DOF; 8 B on the first stage of a step, which
                     needs no dq: 11.2 B on average; dq is rebuilt from the
                     two states, never stored) over the mean stage duration,
                     and the HBM traffic
                     per launch measured with FETCH_SIZE / WRITE_SIZE,
                     calibrated on a plain copy kernel (read from the text
                     summary under profiles/);
  cpu_baseline       the FP64 CPU restatement (OpenMP) timed on this host's
                     cores over one LSRK step of the same brick;
  adapted_mesh       the adapted configuration: a 16^3 forest refined two
                     levels inside a seeded sphere, 2:1 balanced, orders
                     7/5/3 by level, one subdomain per order glued by mortars
                     (hanging faces), FP32, the mortar kernel and the stage
                     kernels of all orders per stage (strategy autotuned);
  serving            a seeded Poisson stream of independent elastic solves
                     (periodic cube bricks of 8..32 elements a side, N in
                     3..7, 200..2000 steps, layered material, Ricker source)
                     admitted under an HBM budget and advanced together, one
                     batched launch per stage, FP32 rates on FP64 state:
                     p50/p99 latency and aggregate DOF-stage-updates/s;
  serving_fp32       the same stream with FP32 state and rates;
  serving_saturated  100 such requests arriving at 1000/s (the GPU completes
                     about 2/s) served in FP32: p50/p99 and the aggregate
                     rate, against the headline and against the same 100
                     requests stepped as one batch (no arrivals, joins or
                     leaves).

--copy-calibration runs only a 1 GiB device-to-device copy kernel (torch)
so that the counters of a kernel of known traffic can be recorded.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per DOF-stage, averaged over the five stages of LSRK4(5).
# FP32 and FP64 keep no dq: every stage reads q and writes the next q, and
# every stage but the first of a step (A = 0) also reads the previous q, from
# which it rebuilds dq: FP32 (8 + 4 x 12) / 5, FP64 (16 + 4 x 24) / 5.
# FP32_RHS (FP64 state, FP32 dq) reads and writes dq as well, except that the
# first stage does not read it and the last does not write it (the next
# step's first stage does not read it): (20 + 3 x 24 + 20) / 5.  Both stage
# kernels move the same bytes.
BYTES_PER_DOF = {"fp32": 11.2, "fp32_rhs": 22.4, "fp64": 22.4, "fp64_plain": 80}
TRAFFIC_FILES = {"node-parallel": "traffic_elastic_stage_fp32_48cube_N7.txt"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def copy_calibration():
    import torch
    n = 1 << 28  # 2^28 float32 = 1 GiB
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log("copy calibration: 3 x 1 GiB device copies done")


def read_traffic(kernel):
    """HBM bytes per launch of the FP32 stage kernel (`kernel`: "line" or
    "node-parallel") from the committed counter summary: FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch) of the kernel, and of a 1 GiB copy kernel,
    which reads and writes exactly 2^30 bytes.  None for a kernel with no
    committed summary."""
    try:
        kv = {}
        with open(os.path.join(ROOT, "profiles", TRAFFIC_FILES[kernel])) as f:
            for line in f:
                if ":" in line and not line.startswith("#"):
                    k, v = line.split(":", 1)
                    kv[k.strip()] = v.strip()
        fk, wk = float(kv["kernel_FETCH_SIZE_KiB"]), float(kv["kernel_WRITE_SIZE_KiB"])
        fc, wc = float(kv["copy_FETCH_SIZE_KiB"]), float(kv["copy_WRITE_SIZE_KiB"])
    except (OSError, KeyError, ValueError):
        return None
    GiB = float(1 << 30)
    return fk * 1024.0 * (GiB / (fc * 1024.0)) + wk * 1024.0 * (GiB / (wc * 1024.0))


def cpu_model():
    """the host CPU's model name (/proc/cpuinfo), or None"""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def time_cpu_steps(cfg, budget_s, max_steps, label):
    """whole LSRK steps of the FP64 C restatement on cfg from rest, until
    budget_s or max_steps: (DOF-stage-updates/s, steps, seconds)"""
    import numpy as np

    import oracle

    m = cfg["mesh"]
    rho, cp, cs = cfg["rho"], cfg["cp"], cfg["cs"]
    lam, mu = rho * (cp * cp - 2 * cs * cs), rho * cs * cs
    P = oracle.OracleProblem(oracle.ELASTIC3D, m.N, m.vmapM, m.vmapP, m.h, rho, lam, mu,
                             x0=m.x0, src=cfg["src"], ops=(m.r, m.w, 1 / m.w, m.Dr))
    q = np.zeros((9, m.K * m.Np))
    ndof = 9 * m.K * m.Np
    steps, el = 0, 0.0
    while steps == 0 or (el < budget_s and steps < max_steps):
        t0 = time.perf_counter()
        q = P.run(q, cfg["dt"], 1, t0=steps * cfg["dt"])
        el += time.perf_counter() - t0
        steps += 1
        log(f"cpu baseline ({label}): {steps} step(s), {el:.1f} s")
    return ndof * 5 * steps / el, steps, el


def run_cpu_baseline(cfg, budget_s):
    """FP64 C restatement (OpenMP) over whole LSRK steps of the same brick on
    every host core, and on one core over a 16^3 brick of the same order and
    material law (a bounded sample: one core would need about half a minute
    per step of the 48^3 brick)."""
    import oracle

    import bfam_amd.workloads as W

    cores = oracle.num_threads()
    rate, steps, el = time_cpu_steps(cfg, budget_s, 4, f"{cores} cores")
    N = cfg["mesh"].N
    small = W.elastic3d_ricker(n=16, N=N, nsteps=200)
    oracle.set_num_threads(1)
    try:
        rate1, steps1, el1 = time_cpu_steps(small, 0.5 * budget_s, 4, "1 core")
    finally:
        oracle.set_num_threads(cores)
    return {"value": rate, "unit": "DOF-stage-updates/s", "cores": cores, "kind": "port",
            "sample": f"{steps} LSRK4(5) step(s) ({5 * steps} stages) of the same 48^3 N={N} "
                      f"elastic brick, FP64, C restatement with OpenMP, {el:.1f} s",
            "cpu_model": cpu_model(),
            "one_core": {"value": rate1, "unit": "DOF-stage-updates/s", "cores": 1,
                         "sample": f"{steps1} LSRK4(5) step(s) of a 16^3 N={N} elastic brick "
                                   f"(same material law and source), FP64, one thread, "
                                   f"{el1:.1f} s"}}


def run_saturated_serving(dev, nreq, budget, headline):
    """nreq requests (8^3..32^3 hexes, N = 3..7, 200..2000 steps) arriving at
    1000/s, far above what one GPU completes (about 2/s), served in FP32
    under the budget; then the same requests stepped as one batch with no
    arrivals, joins or leaves (the rate the serving loop could reach on
    this mix)."""
    import bfam_amd as bx
    reqs = bx.serve_generate(seed=31, nreq=nreq, rate_per_s=1000.0)
    _, st = bx.serve(dev, reqs, bx.FP32, budget_bytes=budget, max_chunk=10)
    subs = [bx.request_build(dev, r, bx.FP32)[0] for r in reqs]
    b = bx.Batch(dev, subs)
    b.autotune(trials=2)
    b.step(1)
    dev.sync()
    ks = 4
    dev.timer_start()
    b.step(ks)
    ms = dev.timer_stop()
    b.close()
    for s_ in subs:
        s_.close()
    dof = sum(9 * r.n ** 3 * (r.N + 1) ** 3 for r in reqs)
    mix = dof * 5 * ks / (ms * 1e-3)
    return {"requests": nreq, "rate_per_s": 1000.0, "budget_bytes": budget,
            "precision": "fp32 state and rates", "p50_s": st["p50_s"], "p99_s": st["p99_s"],
            "max_s": st["max_s"], "dof_stage_per_s": st["dof_stage_per_s"],
            "wall_s": st["wall_s"], "max_active": st["max_active"],
            "batch_rebuilds": st["nbatches"], "batched_mix_dof_stage_per_s": mix,
            "vs_batched_mix": st["dof_stage_per_s"] / mix,
            "vs_headline": st["dof_stage_per_s"] / headline if headline else None}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=48, help="elements per side")
    ap.add_argument("--order", type=int, default=7)
    ap.add_argument("--accuracy-steps", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adapted", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=16)
    ap.add_argument("--serve-saturated", type=int, default=100,
                    help="requests of the saturated serving run (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=0.6, help="arrivals per second")
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--only", choices=["fp32", "fp64", "fp64_plain", "fp32_rhs", "adapted"],
                    help="time one form, or the adapted mesh, only (for profiling runs)")
    ap.add_argument("--eager", action="store_true",
                    help="launch stage by stage instead of replaying a captured graph")
    ap.add_argument("--kernel", choices=["line", "node"],
                    help="run this elastic stage kernel instead of the autotuned one "
                         "(for profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--state", choices=["rest", "random"], default="rest",
                    help="initial state of the timed runs: at rest (the Ricker "
                         "configuration) or seeded random values at every node")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the elastic brick would shard by element slabs with a face-trace exchange "
              "per stage, which is not part of this build)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return

    import numpy as np

    import bfam_amd as bx
    import bfam_amd.workloads as W

    t0 = time.perf_counter()
    cfg = W.elastic3d_ricker(n=args.n, N=args.order, nsteps=200)
    m = cfg["mesh"]
    ndof = 9 * m.K * m.Np
    log(f"setup: {args.n}^3 N={args.order} K={m.K} DOF={ndof} ({time.perf_counter() - t0:.1f} s)")
    if args.state == "random":
        q0 = np.random.default_rng(7).standard_normal((9, m.K * m.Np))
    else:
        q0 = np.zeros((9, m.K * m.Np))
    dev = bx.Device(0)
    K, Wm = args.steps, args.warmup

    def timed(sub, batched, fused):
        sub.set_fields(q0, 0.0, cfg["dt"])
        if fused:
            b = bx.Batch(dev, [sub])
            if args.kernel:
                dev.set_kernel(sub.precision, args.order, 1 if args.kernel == "node" else 0)
            else:
                b.autotune(trials=3)  # stage kernel for this precision and order
            if not batched:
                b.close()  # the single-subdomain path keeps the tuned kernel
        if batched:
            b.capture(1)
            for _ in range(Wm):
                b.launch_graph()
            dev.sync()
            dev.timer_start()
            for _ in range(K):
                b.launch_graph()
            ms = dev.timer_stop()
            b.close()
        else:
            sub.step(Wm)
            dev.sync()
            dev.timer_start()
            sub.step(K)
            ms = dev.timer_stop()
        return ms

    forms = [("fp32", bx.FP32, True), ("fp64", bx.FP64, True),
             ("fp64_plain", bx.FP64_PLAIN, False), ("fp32_rhs", bx.FP32_RHS, True)]
    if args.only:
        forms = [f for f in forms if f[0] == args.only]
    only_adapted = args.only == "adapted"
    res = {}
    for name, prec, batched in forms:
        sub = W.upload(dev, cfg, prec)
        ms = timed(sub, batched and not args.eager, batched)
        sub.close()
        res[name] = {"ms_per_step": ms / K, "ms_per_stage": ms / (5 * K),
                     "dof_stage_per_s": ndof * 5 * K / (ms * 1e-3)}
        if batched:
            res[name]["kernel"] = "node-parallel" if dev.get_kernel(prec, args.order) else "line"
        log(f"{name}: {ms / K:.3f} ms/step, {res[name]['dof_stage_per_s']:.4e} DOF-stage/s")

    accuracy = None
    if not args.only:
        a = W.upload(dev, cfg, bx.FP32)
        b = W.upload(dev, cfg, bx.FP64)
        for s in (a, b):
            s.set_fields(q0, 0.0, cfg["dt"])
            s.step(args.accuracy_steps)
        d2, n2 = a.l2_diff(b)
        accuracy = {"rel_l2_fp32_vs_fp64": float(np.sqrt(d2 / n2)) if n2 > 0 else None,
                    "steps": args.accuracy_steps, "tolerance": 1e-5}
        a.close()
        b.close()
        log(f"accuracy: {accuracy}")

    adapted = None
    if (not args.only and not args.no_adapted) or only_adapted:
        oc = W.elastic3d_octree()
        subs_o = W.upload_multi(dev, oc, bx.FP32)
        for s_, sm in zip(subs_o, oc["subs"]):
            s_.set_fields(np.zeros((9, sm.K * sm.Np)), 0.0, oc["dt"])
        glue = bx.Glue(dev, subs_o, oc["mortars"])
        bo = bx.Batch(dev, subs_o)
        bo.set_glue(glue)
        strategy = bo.autotune(trials=3)
        bo.capture(1)
        for _ in range(3):
            bo.launch_graph()
        dev.sync()
        ka = 40
        dev.timer_start()
        for _ in range(ka):
            bo.launch_graph()
        ms_a = dev.timer_stop()
        adapted = {"elements_per_level": [int(x) for x in oc["K_level"][:3]],
                   "orders": [7, 5, 3], "mortars": int(len(oc["mortars"])),
                   "dof": int(oc["dof"]), "steps": ka, "ms_per_step": ms_a / ka,
                   "dof_stage_per_s": oc["dof"] * 5 * ka / (ms_a * 1e-3),
                   "strategy": "one launch per order" if strategy else "one mixed launch"}
        bo.close()
        glue.close()
        for s_ in subs_o:
            s_.close()
        log(f"adapted mesh: {adapted}")

    serving = {}
    if not args.only and args.serve_requests > 0:
        reqs = bx.serve_generate(seed=2024, nreq=args.serve_requests,
                                 rate_per_s=args.serve_rate)
        budget = int(args.serve_budget_gb * (1 << 30))
        for key, prec, label in (("serving", bx.FP32_RHS, "fp32 rates, fp64 state"),
                                 ("serving_fp32", bx.FP32, "fp32 state and rates")):
            _, st = bx.serve(dev, reqs, prec, budget_bytes=budget, max_chunk=10)
            serving[key] = {
                "requests": args.serve_requests, "rate_per_s": args.serve_rate,
                "budget_bytes": budget, "precision": label,
                "p50_s": st["p50_s"], "p99_s": st["p99_s"], "max_s": st["max_s"],
                "dof_stage_per_s": st["dof_stage_per_s"], "wall_s": st["wall_s"],
                "max_active": st["max_active"], "peak_bytes": st["peak_bytes"],
                "batch_rebuilds": st["nbatches"]}
            log(f"{key}: {serving[key]}")
    if not args.only and args.serve_saturated > 0:
        serving["serving_saturated"] = run_saturated_serving(
            dev, args.serve_saturated, int(args.serve_budget_gb * (1 << 30)),
            res["fp32"]["dof_stage_per_s"] if "fp32" in res else None)
        log(f"serving_saturated: {serving['serving_saturated']}")
    dev.close()
    if only_adapted:
        print(json.dumps({"adapted_mesh": adapted}), flush=True)
        return

    head = res.get("fp32") or next(iter(res.values()))
    head_name = "fp32" if "fp32" in res else next(iter(res))
    launch_s = head["ms_per_stage"] * 1e-3
    head_kernel = head.get("kernel")  # None for the unfused FP64 form
    bpd = BYTES_PER_DOF[head_name]
    achieved = bpd * ndof / launch_s / 1e9
    traffic = read_traffic(head_kernel) if head_name == "fp32" else None

    cpu = None
    if not args.no_cpu and not args.only:
        cpu = run_cpu_baseline(cfg, args.cpu_seconds)

    out = {
        "metric": "DOF-stage-updates/s FP32-batched vs FP64 @relL2<=1e-5; p50/p99 solve latency",
        "value": head["dof_stage_per_s"],
        "unit": "DOF-stage-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wm,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head_name == "fp32" else "f64",
        "data": "synthetic: PCG32 material, Ricker point source from rest"
                if args.state == "rest" else
                "synthetic: PCG32 material, Ricker point source, seeded random initial state",
        "config": {"workload": f"elastic3d_ricker {args.n}^3 periodic hex brick, N={args.order}, "
                               "9 fields, LSRK4(5)", "elements": m.K, "dof": ndof,
                   "form": f"{head_name} fused stage, batched + HIP graph"},
        "forms": res,
        "fp32_over_fp64_plain": (res["fp32"]["dof_stage_per_s"] / res["fp64_plain"]["dof_stage_per_s"]
                                 if "fp32" in res and "fp64_plain" in res else None),
        "accuracy": accuracy,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": (f"elastic_line_kernel<{args.order}, {head_name}>"
                                if head_kernel == "line"
                                else f"elastic_stage_kernel<{args.order}, {head_name}>"
                                if head_kernel else "elastic_part_kernel + plain_update_kernel"),
                     "bytes_per_dof": bpd},
        "cpu_baseline": cpu,
        "adapted_mesh": adapted,
        "serving": serving.get("serving"),
        "serving_fp32": serving.get("serving_fp32"),
        "serving_saturated": serving.get("serving_saturated"),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
