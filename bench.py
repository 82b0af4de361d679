#!/usr/bin/env python3
"""Benchmark of the explicit VMS tet4 time step on one MI355X.

Workload (the element type and size the metric is quoted on): VMS tet4 on a
96^3 x 6 Kuhn-split unit cube (5,308,416 tets, 912,673 nodes), interior node
jitter +-0.15h, parabolic channel inflow u_max = 1, outflow p = 0, no-slip
walls, Re = 200, dt = 0.1 h_min, one dual iteration per step, synthetic
seeded input.  Default precision: FP32 element math / FP64 nodal state; the
plain FP64 path is timed too and reported beside it.

Prints ONE JSON line.  value = element-steps/s of the whole step (all
launches), inputs resident in HBM.  roofline = the dominant kernel's
algorithmic bytes per launch / its mean HIP-event duration on the solver's
stream.  cpu_baseline = the FP64 oracle (plain-C restatement of the
reference's loops, 1 thread) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from femorph_amd import femx, meshgen  # noqa: E402

HBM_PEAK_GBS = 8000.0
F32SUM_NOTE = ("opt-in FEMX_MIXED_F32SUM: FP32 element math with FP32 in-block node sums and FP32 block "
               "partials (sum over blocks, state and update FP64); not the headline, which keeps the residual "
               "accumulation in FP64")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "vms_block_c3_traffic.json")
PROFILED_KERNEL = {"mixed": "vms_block<float, double, true, true, false, false>",
                   "fp64": "vms_block<double, double, true, true, false, false>"}


def profiled_traffic(precision: str, n: int):
    """HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled by the copy calibration),
    or None when the profile does not cover this workload."""
    try:
        t = json.load(open(TRAFFIC_FILE))
    except OSError:
        return None
    if n != 96 or "96^3x6" not in t.get("workload", ""):
        return None
    k = t["kernels"].get(PROFILED_KERNEL[precision])
    return k and k["traffic_bytes"]
METRIC = "element-steps/sec (VMS tet4) + % HBM roofline; p50/p99 solve latency"
FLOPS_FILE = os.path.join(ROOT, "profiles", "supg_block_c2_flops.json")
# vector FMA peak of one MI355X without packed FP32 instructions (this build
# emits none): FP32 then issues at the FP64 rate, 78.6 TFLOP/s for both
VALU_PEAK_TFLOPS = 78.6


def profiled_flops(precision: str, n: int):
    """Flops per launch of the SUPG block kernel at the C2 size, from the
    committed rocprofv3 SQ_INSTS_VALU_* counter runs, or None."""
    try:
        t = json.load(open(FLOPS_FILE))
    except OSError:
        return None
    if n != 128:
        return None
    for k in t["kernels"].values():
        if k["precision"] == precision and k["timing_kind"] == "supg_block_press":
            return k["flops"]
    return None


def algorithmic_bytes_per_element_step(r: float, mixed: bool) -> float:
    """Compulsory HBM bytes per element-step with ideal node reuse, r = N/E:
    conn int32 16 B + subgrid velocity r+w (12 values) + r x (coordinates,
    state u v w p read + write, lumped mass).  FP64: 208 + 96 r; mixed (FP32
    subgrid and coordinates, FP64 state and mass): 112 + 84 r."""
    if mixed:
        return 16 + 96 + r * (12 + 64 + 8)
    return 16 + 192 + r * (24 + 64 + 8)


def make_workload(n=96, jitter=0.15, seed=1):
    m = meshgen.tet_cube(n, jitter, seed)
    bc = meshgen.channel_bc(m, u_max=1.0)
    dt = 0.1 * meshgen.min_tet_h(m)
    params = dict(rho=1.0, mu=1.0 / 200.0, inv_eps=1.0, c1=4.0, c2=2.0, dt=dt)
    ic = meshgen.random_velocity(m, 1e-3, seed)
    return m, bc, params, ic


SETTLE_MAX_S = 4.0


def settle(g, chunk=50, tol=0.01, min_s=0.5, max_s=None):
    """Steps in chunks until two consecutive chunks take the same wall time
    within tol (the GPU clocks and caches have settled; a short window
    straight after start-up runs about 10% slow), for at least min_s and at
    most max_s.  Returns (steps, seconds)."""
    max_s = SETTLE_MAX_S if max_s is None else max_s
    if max_s <= 0:
        return 0, 0.0
    min_s = min(min_s, max_s)
    t_start = time.perf_counter()
    prev, steps = None, 0
    while True:
        t0 = time.perf_counter()
        g.step(chunk)
        g.sync()
        dt = time.perf_counter() - t0
        steps += chunk
        el = time.perf_counter() - t_start
        if (prev is not None and abs(dt - prev) <= tol * prev and el >= min_s) or el >= max_s:
            return steps, el
        prev = dt


def time_solver(dm, bc, params, ic, precision, steps, warmup):
    g = femx.VMSSolver(dm, bc, precision=precision, **params)
    g.set_state(ic)
    g.step(warmup)
    g.sync()
    settled = settle(g)
    t0 = time.perf_counter()
    g.step(steps)
    g.sync()
    t1 = time.perf_counter()
    # per-kernel HIP events on the solver's stream, a second pass of at
    # least 200 steps in the same settled state
    g.kernel_timing(True)
    g.step(max(steps, 200))
    g.sync()
    kt = g.kernel_times()
    g.kernel_timing(False)
    sol, _, _ = g.get_state()
    finite = bool(np.isfinite(sol).all())
    return (t1 - t0), kt, finite, settled


def cpu_baseline(m, bc, params, ic, budget_s=20.0, threads=1):
    """FP64 oracle on whole steps of the same mesh: 1 thread, or the
    element sweeps on `threads` threads (ox_vms_step_threads, bitwise the
    same results; the element-order scatter stays on one thread)."""
    from oracle.oracle import VMSOracle  # test infrastructure, checker only
    o = VMSOracle(m, bc, params["rho"], params["mu"], params["inv_eps"], params["c1"], params["c2"], params["dt"])
    o.set_ic(ic)
    steps, t_tot = 0, 0.0
    while t_tot < budget_s and steps < 10:
        t0 = time.perf_counter()
        o.step(threads=threads)
        t_tot += time.perf_counter() - t0
        steps += 1
    del o
    return m.n_elems * steps / t_tot, steps, t_tot


def host_threads() -> int:
    """Cores this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS when that is set (a shared host's share)."""
    n = len(os.sched_getaffinity(0))
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, n)


def copy_calibration(nbytes: int) -> None:
    """One device-to-device copy of nbytes (HIP's blit kernel): the reference
    point that scales rocprof's FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}), flush=True)


CEILING_GRIDS = (4096, 16384, 65536, 262144)
CEILING_PROBE = ("best of femx_hbm_probe mix_7r5w and mix_7r5w_nt (7 read / 5 write float4 streams, 1 GiB) "
                 "over 4096/16384/65536/262144 workgroups")


def stream_ceiling_gbs(reps: int = 10) -> float:
    """Practical HBM ceiling for a read/write mix like the VMS block kernel's."""
    return max(femx.hbm_probe(k, 1 << 30, g, reps) for k in ("mix_7r5w", "mix_7r5w_nt")
               for g in CEILING_GRIDS) * 1e3


def workload_hbm(args):
    """Streaming bandwidth of float4 grid-stride kernels (femx_hbm_probe):
    the practical ceilings the roofline fractions are read against."""
    res = {}
    for kind in ("copy", "read", "write", "mix_7r5w", "copy_nt", "read_nt", "write_nt", "mix_7r5w_nt"):
        res[kind] = {str(g): round(femx.hbm_probe(kind, 1 << 30, g, 20) * 1e3, 1)
                     for g in (1024, 4096, 16384, 65536, 262144)}
        print(f"hbm {kind}: {res[kind]}", file=sys.stderr, flush=True)
    print(json.dumps({"metric": "HBM streaming bandwidth (float4 grid-stride kernels, 1 GiB per buffer)",
                      "unit": "GB/s", "peak": HBM_PEAK_GBS, "by_kind_and_grid": res,
                      "best": {k: max(v.values()) for k, v in res.items()}}), flush=True)


def _c4_items(seed, precision, count=256):
    """256 lid-driven VMS tet4 cavities, n log-uniform in [8, 48], Re uniform
    in [50, 500], IC uniform +-1e-3 (the C1 physics per mesh)."""
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(np.log(8), np.log(48), count))).astype(int)
    res = rng.uniform(50.0, 500.0, count)
    cache, items, ics = {}, [], []
    for i, (n, re) in enumerate(zip(ns, res)):
        if n not in cache:
            m = meshgen.tet_cube(int(n))
            cache[n] = (m, meshgen.lid_cavity_bc(m), meshgen.min_tet_h(m))
        m, bc, h = cache[n]
        items.append((m, bc, dict(rho=1.0, mu=1.0 / re, inv_eps=1.0, c1=4.0, c2=2.0, dt=0.1 * h, precision=precision)))
        ics.append(meshgen.random_velocity(m, 1e-3, seed + 1000 + i))
    return ns, items, ics


def workload_c4(args):
    """Batch of 256 cavities: one batched launch per phase, graph replay."""
    def note(msg):
        print(f"c4: {msg}", file=sys.stderr, flush=True)

    ns, items, ics = _c4_items(args.seed, femx.FEMX_MIXED)
    note(f"{len(items)} meshes generated")
    E = np.array([m.n_elems for m, _, _ in items], np.float64)
    N = np.array([m.n_nodes for m, _, _ in items], np.float64)
    steps = args.steps
    order = np.argsort(ns, kind="stable")
    sample = [items[j] for j in order[::16]]
    t0 = time.perf_counter()
    be, xcd, tune_ms = femx.vms_batch_autotune(sample, steps=6)
    tune_s = time.perf_counter() - t0
    note(f"autotune: block_elems {be}, xcd_order {xcd}, {tune_s:.1f} s")
    t0 = time.perf_counter()
    b = femx.VMSBatch(items, block_elems=be, xcd_order=xcd)
    build_s = time.perf_counter() - t0
    b.set_state(ics)
    b.step(2)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps, graph=True)
    b.sync()
    secs = time.perf_counter() - t0
    note(f"batch: {steps} steps in {secs:.3f} s")
    info = b.info()
    mixed_states = [b.get_state(i)[0] for i in range(len(items))]
    del b
    # the same batch in FP64: velocity error of the mixed path after 2 + steps steps
    items64 = [(m, bc, dict(p, precision=femx.FEMX_FP64)) for m, bc, p in items]
    b64 = femx.VMSBatch(items64, block_elems=be, xcd_order=xcd)
    b64.set_state(ics)
    b64.step(2 + steps)
    errs = []
    for i in range(len(items)):
        ref = b64.get_state(i)[0]
        errs.append(float(np.linalg.norm(mixed_states[i][:, :3] - ref[:, :3]) / np.linalg.norm(ref[:, :3])))
    del b64
    note("fp64 reference batch done")
    # plain single-mesh path on the same meshes: one handle per mesh, run one after the other
    solvers = []
    for (m, bc, p), ic in zip(items, ics):
        dm = femx.DeviceMesh(m)
        g = femx.VMSSolver(dm, bc, p["rho"], p["mu"], p["inv_eps"], p["c1"], p["c2"], p["dt"], precision=p["precision"])
        g.set_state(ic)
        g.step(2)
        g.sync()
        solvers.append((dm, g))
    t0 = time.perf_counter()
    for _, g in solvers:
        g.step(steps)
        g.sync()
    single_secs = time.perf_counter() - t0
    note(f"single-mesh path: {single_secs:.3f} s")
    del solvers
    total = float(E.sum()) * steps
    value = total / secs
    step_bytes = float(np.sum((112.0 + 84.0 * N / E) * E))
    achieved = step_bytes / (secs / steps) / 1e9
    out = {
        "metric": "element-steps/sec (VMS tet4, batch of 256 cavities)", "value": round(value, 1),
        "unit": "element-steps/s", "n_gpus": 1, "steps": steps, "warmup": 2,
        "ms_per_step": round(secs / steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32-element/f64-state", "data": "synthetic (seeded cavities)",
        "config": {"workload": f"C4: 256 VMS tet4 lid cavities, n log-uniform [8,48], Re U[50,500] "
                               f"({int(E.sum())} tets, {int(N.sum())} nodes)",
                   "elements": int(E.sum()), "nodes": int(N.sum()), "meshes": len(items),
                   "n_min": int(ns.min()), "n_max": int(ns.max())},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole batched step (block kernel + fix-up + time tick)",
                     "step_bytes": step_bytes},
        "batch": dict(info, graph=True, autotune={"block_elems": be, "xcd_order": xcd,
                                                  "ms_per_step": [round(x, 4) for x in tune_ms],
                                                  "candidates": "256/224/192 x xcd 1/0", "sample_meshes": len(sample),
                                                  "seconds": round(tune_s, 2)},
                      build_seconds=round(build_s, 2)),
        "single_mesh_path": {"value": round(total / single_secs, 1), "ms_per_step": round(single_secs / steps * 1e3, 4),
                             "batch_speedup": round(single_secs / secs, 2)},
        "error_vs_fp64": {"steps": 2 + steps, "max_rel_l2_velocity": max(errs),
                          "median_rel_l2_velocity": float(np.median(errs))},
    }
    if not args.no_cpu:
        from oracle.oracle import VMSOracle  # test infrastructure, checker only
        j = int(order[len(order) // 2])
        m, bc, p = items[j]
        o = VMSOracle(m, bc, p["rho"], p["mu"], p["inv_eps"], p["c1"], p["c2"], p["dt"])
        o.set_ic(ics[j])
        k, tt = 0, 0.0
        while tt < args.cpu_budget and k < 200:
            t0 = time.perf_counter()
            o.step()
            tt += time.perf_counter() - t0
            k += 1
        cv = m.n_elems * k / tt
        out["cpu_baseline"] = {"value": round(cv, 1), "unit": "element-steps/s", "cores": 1, "kind": "port",
                               "sample": f"{k} FP64 steps of the median mesh ({m.n_elems} tets), oracle, 1 thread"}
        out["speedup_vs_cpu_1t"] = round(value / cv, 1)
    print(json.dumps(out), flush=True)


def workload_serve(args):
    """Poisson arrivals of tet4 (VMS) / hexa8 (SUPG) cavity solves, n
    log-uniform in [8, 64], 50..500 steps, mixed precision; per rate the
    latency percentiles and throughput."""
    rng = np.random.default_rng(args.seed)
    cache = {}

    def mesh_for(kind, n):
        if (kind, n) not in cache:
            if kind == "vms":
                m = meshgen.tet_cube(n, 0.15, n)
                cache[(kind, n)] = (m, meshgen.lid_cavity_bc(m), 0.1 * meshgen.min_tet_h(m))
            else:
                m = meshgen.hex_cube(n, 0.2, n)
                cache[(kind, n)] = (m, meshgen.lid_cavity_bc(m), 0.1 / n)
        return cache[(kind, n)]

    rates = [float(r) for r in args.rates.split(",")]
    plans = []
    for rate in rates:
        arr = np.cumsum(rng.exponential(1.0 / rate, args.requests))
        reqs = []
        for _ in range(args.requests):
            kind = "vms" if rng.random() < 0.5 else "supg"
            if args.serve_kinds != "both":
                kind = args.serve_kinds
            n = int(np.rint(np.exp(rng.uniform(np.log(8), np.log(64)))))
            steps = int(rng.integers(50, 501))
            re = float(rng.uniform(50.0, 500.0))
            mesh_for(kind, n)
            reqs.append((kind, n, steps, re))
        plans.append((rate, arr, reqs))
    # one untimed request of each kind first: the process's first solves
    # load the kernels' code and fill the device memory cache
    srv = femx.Server(budget_bytes=args.budget_gb << 30, max_active=args.max_active, prep_threads=args.prep_threads,
                      chunk_steps=args.chunk, gpu_threads=args.gpu_threads)
    for kind in ("vms", "supg"):
        if args.serve_kinds in ("both", kind):
            m, bc, dt = mesh_for(kind, 16)
            if kind == "vms":
                srv.submit_vms(m, bc, 60, 1.0, 1.0 / 200, 1.0, 4.0, 2.0, dt)
            else:
                srv.submit_supg(m, bc, 60, 1.0, 1.0 / 200, 1.0, 4.0, dt)
    srv.drain(120.0)
    for s_ in srv.poll():
        srv.result(s_["id"])
    srv.close()
    table = {}
    for rate, arr, reqs in plans:
        srv = femx.Server(budget_bytes=args.budget_gb << 30, max_active=args.max_active, prep_threads=args.prep_threads,
                          chunk_steps=args.chunk, gpu_threads=args.gpu_threads)
        ids = {}
        t0 = time.perf_counter()
        for t_arr, (kind, n, steps, re) in zip(arr, reqs):
            wait = t0 + t_arr - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            m, bc, dt = mesh_for(kind, n)
            if kind == "vms":
                rid = srv.submit_vms(m, bc, steps, 1.0, 1.0 / re, 1.0, 4.0, 2.0, dt)
            else:
                rid = srv.submit_supg(m, bc, steps, 1.0, 1.0 / re, 1.0, 4.0, dt)
            ids[rid] = m.n_elems * steps
        srv.drain(1200.0)
        st = srv.poll()
        for s_ in st:
            srv.result(s_["id"])
        srv.close()
        lat = np.array([s_["t_done"] - s_["t_submit"] for s_ in st])
        wait = np.array([s_["t_admit"] - s_["t_submit"] for s_ in st])
        prep = np.array([s_["t_start"] - s_["t_admit"] for s_ in st])
        run = np.array([s_["t_done"] - s_["t_start"] for s_ in st])
        span = max(s_["t_done"] for s_ in st) - min(s_["t_submit"] for s_ in st)
        done = [s_ for s_ in st if s_["status"] in (0, femx.EXIT_CONDITION)]
        es = sum(s_["elements"] * s_["steps_done"] for s_ in done)
        # steady state: completions between the 10th and 90th percentile of
        # the completion times, over that window (no ramp-up, no drain)
        td = np.sort([s_["t_done"] for s_ in st])
        lo, hi = td[len(td) // 10], td[(9 * len(td)) // 10]
        steady = ((td > lo) & (td <= hi)).sum() / (hi - lo) if hi > lo else float("nan")
        table[str(int(rate))] = {"requests": len(st), "ok": len(done), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
                                 "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
                                 "req_per_s": round(len(st) / span, 2), "steady_req_per_s": round(float(steady), 2),
                                 "element_steps_per_s": round(es / span, 1),
                                 "offered_req_per_s": round(len(st) / float(arr[-1]), 2),
                                 "mean_ms": {"queue": round(float(wait.mean()) * 1e3, 2),
                                             "prep": round(float(prep.mean()) * 1e3, 2),
                                             "resident": round(float(run.mean()) * 1e3, 2)}}
        by_kind = {}
        for kname, kid in (("vms", 0), ("supg", 1)):
            sel = np.array([s_["kind"] == kid for s_ in st])
            if sel.any():
                by_kind[kname] = {"requests": int(sel.sum()),
                                  "p50_ms": round(float(np.percentile(lat[sel], 50)) * 1e3, 2),
                                  "p99_ms": round(float(np.percentile(lat[sel], 99)) * 1e3, 2),
                                  "max_prep_ms": round(float(prep[sel].max()) * 1e3, 2),
                                  "mean_prep_ms": round(float(prep[sel].mean()) * 1e3, 2),
                                  "mean_resident_ms": round(float(run[sel].mean()) * 1e3, 2)}
        table[str(int(rate))]["by_kind"] = by_kind
        print(json.dumps({"rate": rate, **table[str(int(rate))]}), file=sys.stderr, flush=True)
    top = table[str(int(rates[-1]))]
    out = {"metric": "p50/p99 solve latency (served VMS tet4 / SUPG hexa8 requests)",
           "value": top["element_steps_per_s"], "unit": "element-steps/s", "n_gpus": 1, "steps": None, "warmup": 0,
           "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32-element/f64-state", "data": "synthetic (seeded jittered cavities)",
           "config": {"workload": f"C5: Poisson arrivals at {args.rates} req/s, {args.requests} requests per rate, "
                                  "tet4 (VMS) or hexa8 (SUPG) cavities, n log-uniform [8,64], steps U[50,500]",
                      "budget_gb": args.budget_gb, "max_active": args.max_active, "chunk_steps": args.chunk,
                      "prep_threads": args.prep_threads, "gpu_threads": args.gpu_threads,
                      "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "runtime default"),
                      "warmup": "one untimed 16^3 cavity request of each kind, 60 steps, on a server of its own"},
           "latency": table}
    print(json.dumps(out), flush=True)


def workload_c1(args):
    """C1: VMS tet4 20^3 x 6 lid cavity, Re 100 (48,000 tets).  It moves
    about 6 MB per step and its 188 blocks fill a small part of the chip,
    so its step is bound by the latency of one block, not by HBM: the
    single handle's launches, and the same mesh as a batch of one replayed
    from a HIP graph.  CPU: the oracle over all 100 steps."""
    m = meshgen.tet_cube(20)
    bc = meshgen.lid_cavity_bc(m)
    p = dict(rho=1.0, mu=0.01, inv_eps=1.0, c1=4.0, c2=2.0, dt=0.1 * meshgen.min_tet_h(m))
    ic = meshgen.random_velocity(m, 1e-3, args.seed)
    E = m.n_elems
    k = args.steps
    dm = femx.DeviceMesh(m)
    res = {}
    for name, prec in (("mixed", femx.FEMX_MIXED), ("fp64", femx.FEMX_FP64)):
        g = femx.VMSSolver(dm, bc, precision=prec, **p)
        g.set_state(ic)
        g.step(max(args.warmup, 100))
        g.sync()
        t0 = time.perf_counter()
        g.step(k)
        g.sync()
        secs = time.perf_counter() - t0
        g.kernel_timing(True)
        g.step(k)
        g.sync()
        kt = g.kernel_times()
        g.kernel_timing(False)
        res[name] = {"value": round(E * k / secs, 1), "ms_per_step": round(secs / k * 1e3, 5),
                     "kernels_us": {n_: round(v["total_ms"] / v["launches"] * 1e3, 2)
                                    for n_, v in kt.items() if v["launches"]}}
        del g
    b = femx.VMSBatch([(m, bc, dict(p, precision=femx.FEMX_MIXED))])
    b.set_state([ic])
    b.step(max(args.warmup, 100))
    b.sync()
    t0 = time.perf_counter()
    b.step(k)
    b.sync()
    bsecs = time.perf_counter() - t0
    del b
    main_ = res["mixed"]
    step_bytes = (112.0 + 84.0 * m.n_nodes / E) * E
    out = {"metric": "element-steps/sec (VMS tet4)", "value": main_["value"], "unit": "element-steps/s",
           "n_gpus": 1, "steps": k, "warmup": max(args.warmup, 100), "ms_per_step": main_["ms_per_step"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32-element/f64-state",
           "data": "synthetic (seeded Kuhn-split cube, lid cavity)",
           "config": {"workload": f"C1: VMS tet4 20^3x6 lid cavity Re100 ({E} tets, {m.n_nodes} nodes)",
                      "elements": E, "nodes": m.n_nodes},
           "roofline": {"bound": "latency of one block", "note": "the step's compulsory bytes "
                        f"({step_bytes / 1e6:.1f} MB) take {step_bytes / (HBM_PEAK_GBS * 1e9) * 1e6:.2f} us at "
                        "the HBM peak; not judged against HBM"},
           "fp64": res["fp64"], "mixed_kernels_us": main_["kernels_us"],
           "batch_of_one_graph": {"value": round(E * k / bsecs, 1), "ms_per_step": round(bsecs / k * 1e3, 5)}}
    if not args.no_cpu:
        from oracle.oracle import VMSOracle  # test infrastructure, checker only
        for threads in sorted({1, host_threads()}):
            o = VMSOracle(m, bc, p["rho"], p["mu"], p["inv_eps"], p["c1"], p["c2"], p["dt"])
            o.set_ic(ic)
            t0 = time.perf_counter()
            o.step(100, threads=threads)
            cs = time.perf_counter() - t0
            del o
            key = "cpu_baseline" if threads == 1 else "cpu_baseline_all_cores"
            out[key] = {"value": round(E * 100 / cs, 1), "unit": "element-steps/s", "cores": threads, "kind": "port",
                        "sample": f"all 100 FP64 steps of C1, oracle ({threads} thread{'s' if threads > 1 else ''}), "
                                  f"{cs:.2f} s"}
    print(json.dumps(out), flush=True)


def workload_union(args):
    """Unions of resident solves: 10 and 40 cavities of the serving size mix
    (n log-uniform in [8, 64], seeded), VMS tet4 and SUPG hexa8 in mixed
    precision, stepped together by femx_{vms,supg}_union_step against the
    same handles each stepped on its own stream."""
    rng = np.random.default_rng(args.seed + 6)
    L = femx.load()
    rows = []
    for kind in ("vms", "supg"):
        cache = {}
        for count in (10, 40):
            ns = np.rint(np.exp(rng.uniform(np.log(8), np.log(64), count))).astype(int)
            hs, E, skipped = [], 0, []
            for n in ns:
                n = int(n)
                if n not in cache:
                    if kind == "vms":
                        m = meshgen.tet_cube(n, 0.15, n)
                        cache[n] = (m, meshgen.lid_cavity_bc(m), 0.1 * meshgen.min_tet_h(m))
                    else:
                        m = meshgen.hex_cube(n, 0.2, n)
                        cache[n] = (m, meshgen.lid_cavity_bc(m), 0.1 / n)
                m, bc, dt = cache[n]
                dm = femx.DeviceMesh(m)
                if kind == "vms":
                    g = femx.VMSSolver(dm, bc, 1.0, 1.0 / 200, 1.0, 4.0, 2.0, dt, precision=femx.FEMX_MIXED)
                    g.set_state(np.zeros((m.n_nodes, 4)))
                    g.step(1)
                else:
                    g = femx.SUPGSolver(dm, bc, 1.0, 1.0 / 200, 1.0, 4.0, dt, precision=femx.FEMX_MIXED)
                    g.set_state(np.zeros((m.n_nodes, 3)))
                    g.step(1)
                    if not L.femx_supg_union_accepts(g._h):
                        skipped.append(n)
                        continue
                hs.append((dm, g))
                E += m.n_elems
            u = femx.VMSUnion() if kind == "vms" else femx.SUPGUnion()
            for dm, g in hs:
                u.add(g, 1 << 30)
            k = args.steps
            u.step(max(k // 10, 2))
            u.sync()
            t0 = time.perf_counter()
            u.step(k)
            u.sync()
            t1 = time.perf_counter()
            for dm, g in hs:
                u.remove(g)
            if kind == "supg":
                for dm, g in hs:
                    g.finish()
            # the same members, each on its own stream
            ks = max(k // 4, 2)
            for rep_ in range(2):
                ta = time.perf_counter()
                for dm, g in hs:
                    if kind == "vms":
                        g.step(ks if rep_ else 2)
                    else:
                        g.step_async(ks if rep_ else 2)
                for dm, g in hs:
                    g.sync() if kind == "vms" else g.finish()
                tb = time.perf_counter()
            row = {"kind": kind, "members": len(hs), "elements": int(E), "skipped_sizes": skipped,
                   "union_ms_per_step": round((t1 - t0) / k * 1e3, 4),
                   "union_element_steps_per_s": round(E * k / (t1 - t0), 1),
                   "own_streams_element_steps_per_s": round(E * ks / (tb - ta), 1)}
            print(json.dumps(row), file=sys.stderr, flush=True)
            rows.append(row)
            del u, hs
    top = rows[1]
    out = {"metric": "element-steps/s of a union of resident solves", "value": top["union_element_steps_per_s"],
           "unit": "element-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": max(args.steps // 10, 2),
           "ms_per_step": top["union_ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32-element/f64-state", "data": "synthetic (seeded jittered cavities)",
           "config": {"workload": "unions of 10 and 40 resident lid cavities, n log-uniform [8,64]: VMS tet4 "
                                  "and SUPG hexa8 (value: 40 VMS members)"},
           "unions": rows}
    print(json.dumps(out), flush=True)


def workload_c2(args):
    """SUPG hexa8 on a jittered 128^3 cube (2,097,152 hexa, 2,146,689 nodes),
    Re 400 lid cavity, AB2 + B-hat + K1..K4, mixed and FP64."""
    n = args.n if args.n != 96 else 128
    m = meshgen.hex_cube(n, 0.2, args.seed)
    bc = meshgen.lid_cavity_bc(m)
    dm = femx.DeviceMesh(m)
    p = dict(rho=1.0, mu=1.0 / 400.0, alpha=1.0, cI=4.0, dt=0.1 / n)
    ic = meshgen.random_velocity(m, 1e-3, args.seed)[:, :3]
    E, N = m.n_elems, m.n_nodes
    res = {}
    wanted = args.precisions.split(",")
    for name, prec in (("mixed", femx.FEMX_MIXED), ("fp64", femx.FEMX_FP64), ("mixed_f32sum", femx.FEMX_MIXED_F32SUM)):
        if name not in wanted:
            continue
        g = femx.SUPGSolver(dm, bc, precision=prec, **p)
        g.set_state(ic)
        g.step(args.warmup)
        g.sync()
        settle(g)
        t0 = time.perf_counter()
        g.step(args.steps)
        g.sync()
        secs = time.perf_counter() - t0
        g.kernel_timing(True)
        g.step(args.steps)
        g.sync()
        kt = g.kernel_times()
        g.kernel_timing(False)
        res[name] = (secs, kt)
        print(f"c2: {name} {secs / args.steps * 1e3:.4f} ms/step", file=sys.stderr, flush=True)
    secs, kt = res[args.precision]
    dom_name = max(kt, key=lambda k: kt[k]["total_ms"])
    dom = kt[dom_name]
    dom_ms = dom["total_ms"] / max(dom["launches"], 1)
    achieved = dom["bytes_per_launch"] / (dom_ms * 1e-3) / 1e9
    flops = profiled_flops(args.precision, n)
    tflops = flops / (dom_ms * 1e-3) / 1e12 if flops else None
    valu_probe = femx.valu_probe("f32" if args.precision == "mixed" else "f64", 32768, 2048, 5)
    other = "fp64" if args.precision == "mixed" else "mixed"
    out = {"metric": "element-steps/sec (SUPG hexa8)", "value": round(E * args.steps / secs, 1),
           "unit": "element-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(secs / args.steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32-element/f64-state" if args.precision == "mixed" else "f64",
           "data": "synthetic (seeded jittered cube)",
           "config": {"workload": f"C2: SUPG hexa8 {n}^3 jitter0.2 lid cavity Re400, AB2+Bhat+K1-K4 "
                                  f"({E} hexa, {N} nodes)", "elements": E, "nodes": N},
           "roofline": {"bound": "valu", "achieved": round(tflops, 2) if tflops else None,
                        "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(tflops / VALU_PEAK_TFLOPS, 4) if tflops else None, "traffic": None,
                        "kernel": dom_name, "kernel_ms": round(dom_ms, 4), "flops_per_launch": flops,
                        "flops_source": "profiles/supg_block_c2_flops.json (rocprofv3 SQ_INSTS_VALU_* counters)",
                        "peak_note": "vector FMA peak without packed FP32 (FP32 then issues at the FP64 rate)",
                        "valu_probe_tflops": round(valu_probe, 2),
                        "valu_probe_frac": round(tflops / valu_probe, 4) if tflops else None,
                        "hbm": {"achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": round(achieved / HBM_PEAK_GBS, 4), "kernel_bytes": dom["bytes_per_launch"]},
                        "kernels": {k: dict(launches=v["launches"], mean_ms=round(v["total_ms"] / max(v["launches"], 1), 4))
                                    for k, v in kt.items()}},
           }
    for name in (other, "mixed_f32sum"):
        if name in res:
            out[name] = {"value": round(E * args.steps / res[name][0], 1),
                         "ms_per_step": round(res[name][0] / args.steps * 1e3, 4)}
    if "mixed_f32sum" in out:
        out["mixed_f32sum"]["note"] = F32SUM_NOTE
    print(json.dumps(out), flush=True)


def make_ins_workload(n, seed, hexa=False):
    """Implicit advection-diffusion (femIncompressibleSolver, STAGES 0 1,
    PRESCRIBEDVEL 0): a Gaussian blob in the time-periodic vortex on a
    jittered n^3 x 6 tet cube (or n^3 hexa cube), zero Dirichlet walls,
    diffusivity 1e-3, dt = 0.5 h, generalized-alpha with rho_inf = 0.5."""
    m = meshgen.hex_cube(n, 0.2, seed) if hexa else meshgen.tet_cube(n, 0.15, seed)
    c = m.coords
    wall = np.where((c.min(axis=1) < 1e-12) | (c.max(axis=1) > 1 - 1e-12))[0].astype(np.int32)
    phi = np.exp(-40.0 * ((c[:, 0] - 0.5) ** 2 + (c[:, 1] - 0.3) ** 2 + (c[:, 2] - 0.5) ** 2))
    phi[wall] = 0.0
    rinf = 0.5
    am = 0.5 * (3.0 - rinf) / (1.0 + rinf)
    af = 1.0 / (1.0 + rinf)
    params = dict(dt=0.5 / n, alpha_m=am, alpha_f=af, gamma=0.5 + am - af)
    diff = np.full((m.n_elems, 3), 1.0e-3 / np.sqrt(3.0))
    return m, wall, phi, diff, params


def workload_ins(args):
    """Implicit transient advection-diffusion stage on the GPU: per step the
    prescribed vortex velocity, assembly, Dirichlet rows, GMRES to the
    reference's 1e-5, generalized-alpha update."""
    n = args.n
    hexa = args.ins_elem == "hexa8"
    m, wall, phi0, diff, p = make_ins_workload(n, args.seed, hexa)
    shape = f"hexa8 {n}^3" if hexa else f"tet4 {n}^3x6"
    E, N = m.n_elems, m.n_nodes
    dm = femx.DeviceMesh(m)
    t0 = time.perf_counter()
    g = femx.INSSolver(dm, el_diff=diff, dir_node=wall, dir_val=np.zeros((len(wall), 3)), **p)
    setup_s = time.perf_counter() - t0
    g.set_state(phi=phi0, phi_dot=np.zeros(N))
    dt = p["dt"]

    def run(k, t):
        its = []
        for _ in range(k):
            g.prescribe_velocity(0, t)
            its.append(g.advdiff_solve()["iterations"])
            g.update()
            t += dt
        g.sync()
        return t, its

    t, _ = run(args.warmup, 0.0)
    t0 = time.perf_counter()
    t, its = run(args.steps, t)
    secs = time.perf_counter() - t0
    g.kernel_timing(True)
    t, its2 = run(args.steps, t)
    kt = g.kernel_times()
    g.kernel_timing(False)
    phi, _, _ = g.get_state()
    _, col, _, _, _ = g.get_system()
    nnz = len(col)
    value = E * args.steps / secs
    dom = max(kt, key=lambda k: kt[k]["total_ms"])
    d = kt[dom]
    mean_ms = d["total_ms"] / max(d["launches"], 1)
    achieved = d["bytes_per_launch"] / (mean_ms * 1e-3) / 1e9
    out = {
        "metric": f"element-steps/sec (implicit advection-diffusion, {shape.split()[0]})", "value": round(value, 1),
        "unit": "element-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(secs / args.steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded jittered Kuhn-split cube)",
        "config": {"workload": f"implicit advection-diffusion {shape}, vortex velocity, Gaussian blob, "
                               f"GMRES(30) to 1e-5 ({E} elements, {N} nodes, {nnz} non-zeros)",
                   "elements": E, "nodes": N, "nnz": nnz, **{k: round(v, 6) for k, v in p.items()}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
                     "kernel_ms": round(mean_ms, 4), "kernel_bytes": d["bytes_per_launch"],
                     "kernels": {k: {"launches": v["launches"], "mean_ms": round(v["total_ms"] / max(v["launches"], 1), 4),
                                     "bytes": round(v["bytes_per_launch"]),
                                     "gbs": round(v["bytes_per_launch"] / (v["total_ms"] / max(v["launches"], 1) * 1e-3) / 1e9, 1)
                                     if v["launches"] else None,
                                     "ms_per_step": round(v["total_ms"] / args.steps, 4)} for k, v in kt.items()}},
        "gmres_iterations_per_step": round(float(np.mean(its + its2)), 2),
        "setup_seconds": round(setup_s, 2),
        "finite_state": bool(np.isfinite(phi).all()),
    }
    if not args.no_cpu:
        from oracle.oracle import INSOracle  # test infrastructure, checker only
        o = INSOracle(m, diff, wall, np.zeros((len(wall), 3)), p["dt"], p["alpha_m"], p["alpha_f"], p["gamma"])
        o.phi = phi0.copy()
        k, tt, tc = 0, 0.0, 0.0
        while tt < args.cpu_budget and k < 10:
            t0 = time.perf_counter()
            o.prescribe(0, tc)
            o.advdiff_solve()
            o.update()
            tt += time.perf_counter() - t0
            tc += dt
            k += 1
        cv = E * k / tt
        out["cpu_baseline"] = {"value": round(cv, 1), "unit": "element-steps/s", "cores": 1, "kind": "port",
                               "sample": f"{k} full steps of the same mesh (assembly + GMRES(30) to 1e-5 + update), "
                                         f"oracle (plain C, 1 thread), {tt:.1f} s"}
        out["speedup_vs_cpu_1t"] = round(value / cv, 1)
    print(json.dumps(out), flush=True)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def workload_dropin(args):
    """femVMSExplicitFluidSolver::solve as shipped: femx_explicit --vms
    --mixed on a model.txt of the C3 workload (zero initial state, no saves
    inside the window), against the C-ABI step on the same mesh.  The
    drop-in prints one line per dual iteration (per-iteration norms and CFL,
    from the device ring of the last 64 steps); the per-step rate comes from
    the arrival times of those lines on its stdout pipe, from the line of
    step steps/3 to the last one, so the model read, the set-up and the
    clocks' ramp after the read are not in the window."""
    import subprocess
    import tempfile
    from femorph_amd import modelfile
    from femorph_amd.build import DRIVER

    m, bc, params, ic = make_workload(args.n)
    E = m.n_elems
    steps = max(args.steps, 3000)
    d = tempfile.mkdtemp(prefix="femx_dropin_")
    path = os.path.join(d, "model.txt")
    t0 = time.perf_counter()
    modelfile.write_model_txt(path, m, bc, time_step=params["dt"], total_steps=steps + 1, save_every=10 ** 9,
                              vmsprops=(params["rho"], params["mu"], params["inv_eps"], params["c1"], params["c2"]))
    print(f"dropin: model.txt written in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    # the C-ABI step on the same mesh: no diagnostics, and the drop-in's
    # norms-only diagnostics, both on the handle's stream; timed first, so
    # that the drop-in process then starts on a GPU whose clocks have settled
    dm = femx.DeviceMesh(m)
    res = {}
    for name, diag in (("c_abi", 0), ("c_abi_norms", 2)):
        g = femx.VMSSolver(dm, bc, precision=femx.FEMX_MIXED, **params)
        g.set_state(np.zeros((m.n_nodes, 4)))
        g.set_diagnostics(diag)
        g.step(args.warmup)
        g.sync()
        settle(g)
        t0 = time.perf_counter()
        g.step(steps)
        g.sync()
        res[name] = (time.perf_counter() - t0) / steps * 1e3
        del g
    del dm
    t_start = time.perf_counter()
    pr = subprocess.Popen([DRIVER, "--vms", "--mixed", "model.txt"], cwd=d, stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE)
    marks, lines, tail = [], 0, b""
    while True:
        chunk = pr.stdout.read1(1 << 16)
        if not chunk:
            break
        data = tail + chunk
        lines += data.count(b"\n")
        tail = data[data.rfind(b"\n") + 1:]
        marks.append((time.perf_counter(), lines))
    rc = pr.wait()
    wall = time.perf_counter() - t_start
    err = pr.stderr.read().decode(errors="replace")
    for f in os.listdir(d):
        os.remove(os.path.join(d, f))
    os.rmdir(d)
    if rc != 0 or lines < steps:
        raise RuntimeError(f"femx_explicit failed (rc {rc}, {lines} lines): {err[-2000:]}")
    a = next(mk for mk in marks if mk[1] >= steps // 3)
    b = marks[-1]
    dropin_ms = (b[0] - a[0]) / (b[1] - a[1]) * 1e3
    print(f"dropin: {lines} lines, {dropin_ms:.4f} ms/step, whole run {wall:.1f} s", file=sys.stderr, flush=True)
    out = {"metric": "ms per VMS tet4 step through femVMSExplicitFluidSolver::solve (drop-in) vs the C ABI",
           "value": round(dropin_ms, 4), "unit": "ms/step", "higher_is_better": False,
           "dtype": "f32-element/f64-state", "data": "synthetic (seeded jittered Kuhn-split cube, channel BCs)",
           "config": {"workload": f"drop-in solve() on VMS tet4 {args.n}^3x6 channel Re200 ({E} tets), "
                                  f"{steps} steps, zero initial state", "elements": E, "steps": steps},
           "element_steps_per_s": round(E / (dropin_ms * 1e-3), 1),
           "c_abi_ms_per_step": round(res["c_abi"], 4),
           "c_abi_with_norms_ms_per_step": round(res["c_abi_norms"], 4),
           "dropin_over_c_abi": round(dropin_ms / res["c_abi"], 3),
           "whole_run_s": round(wall, 2)}
    print(json.dumps(out), flush=True)


def workload_accuracy(args):
    """Error of the FP32-element / FP64-state path after `steps` steps:
    C1 against the FP64 oracle itself, C3 (VMS) and C2 (SUPG) against the
    plain FP64 GPU path (pinned to the oracle at <= 1e-12 per step)."""
    steps = args.steps
    out = {"metric": "relative L2 error of the mixed path after N steps", "steps": steps, "configs": {}}

    def note(msg):
        print(f"accuracy: {msg}", file=sys.stderr, flush=True)

    # C1: 20^3 x 6 lid cavity, Re 100, vs the oracle
    from oracle.oracle import VMSOracle  # the checker
    m = meshgen.tet_cube(20)
    bc = meshgen.lid_cavity_bc(m)
    p = dict(rho=1.0, mu=0.01, inv_eps=1.0, c1=4.0, c2=2.0, dt=0.1 * meshgen.min_tet_h(m))
    ic = meshgen.random_velocity(m, 1e-3, args.seed)
    o = VMSOracle(m, bc, p["rho"], p["mu"], p["inv_eps"], p["c1"], p["c2"], p["dt"])
    o.set_ic(ic)
    dm = femx.DeviceMesh(m)
    res = {}
    names = {femx.FEMX_MIXED: "mixed", femx.FEMX_FP64: "fp64", femx.FEMX_MIXED_F32SUM: "mixed_f32sum"}
    for prec in names:
        g = femx.VMSSolver(dm, bc, precision=prec, **p)
        g.set_state(ic)
        res[prec] = g
    done = 0
    for chunk in (100,) * (steps // 100) + ((steps % 100,) if steps % 100 else ()):
        o.step(chunk)
        for g in res.values():
            g.step(chunk)
        done += chunk
        note(f"C1 {done} steps")
    ref = o.get("sol_n")
    c1 = {}
    for prec, g in res.items():
        sol = g.get_state()[0]
        c1[names[prec]] = {
            "rel_l2_velocity": _rel(sol[:, :3], ref[:, :3]), "rel_l2_pressure": _rel(sol[:, 3], ref[:, 3])}
    out["configs"]["C1 VMS tet4 20^3x6 cavity Re100 (vs FP64 oracle)"] = c1
    del res, dm, o
    # C3: 96^3 x 6 channel Re 200, mixed vs FP64 GPU path
    m, bc, params, ic = make_workload(96)
    dm = femx.DeviceMesh(m)
    gm = femx.VMSSolver(dm, bc, precision=femx.FEMX_MIXED, **params)
    gd = femx.VMSSolver(dm, bc, precision=femx.FEMX_FP64, **params)
    gs = femx.VMSSolver(dm, bc, precision=femx.FEMX_MIXED_F32SUM, **params)
    for g in (gm, gd, gs):
        g.set_state(ic)
        g.step(steps)
    a, b, c = gm.get_state()[0], gd.get_state()[0], gs.get_state()[0]
    out["configs"]["C3 VMS tet4 96^3x6 channel Re200 (vs FP64 GPU path)"] = {
        "rel_l2_velocity": _rel(a[:, :3], b[:, :3]), "rel_l2_pressure": _rel(a[:, 3], b[:, 3]),
        "mixed_f32sum": {"rel_l2_velocity": _rel(c[:, :3], b[:, :3]), "rel_l2_pressure": _rel(c[:, 3], b[:, 3])}}
    note("C3 done")
    del gm, gd, gs, dm
    # C2: SUPG hexa8 128^3 jittered lid cavity Re 400, AB2 + B-hat + K1..K4
    mh = meshgen.hex_cube(128, 0.2, args.seed)
    bch = meshgen.lid_cavity_bc(mh)
    dmh = femx.DeviceMesh(mh)
    ph = dict(rho=1.0, mu=1.0 / 400.0, alpha=1.0, cI=4.0, dt=0.1 / 128)
    ich = meshgen.random_velocity(mh, 1e-3, args.seed)[:, :3]
    sm = femx.SUPGSolver(dmh, bch, precision=femx.FEMX_MIXED, **ph)
    sd = femx.SUPGSolver(dmh, bch, precision=femx.FEMX_FP64, **ph)
    ss = femx.SUPGSolver(dmh, bch, precision=femx.FEMX_MIXED_F32SUM, **ph)
    for g in (sm, sd, ss):
        g.set_state(ich)
        g.step(steps)
    a, b, c = sm.get_state(), sd.get_state(), ss.get_state()
    out["configs"]["C2 SUPG hexa8 128^3 jitter0.2 cavity Re400 (vs FP64 GPU path)"] = {
        "rel_l2_velocity": _rel(a["sol"], b["sol"]), "rel_l2_pressure": _rel(a["pres"], b["pres"]),
        "mixed_f32sum": {"rel_l2_velocity": _rel(c["sol"], b["sol"]), "rel_l2_pressure": _rel(c["pres"], b["pres"])}}
    note("C2 done")
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default per workload: long enough that the start-up of a timed "
                         "window, about 1 ms, stays below 1%% of it)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=96, help="cube edge count (96 -> 5.31M tets)")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (profiler counter calibration)")
    ap.add_argument("--workload", choices=["c3", "c1", "c2", "c4", "serve", "union", "accuracy", "ins", "hbm",
                                           "dropin"],
                    default="c3",
                    help="c3: one 5.31M-tet VMS mesh (default); c4: batch of 256 cavities; serve: request "
                         "stream; hbm: streaming-bandwidth probe")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--precisions", default="mixed,fp64,mixed_f32sum",
                    help="c3/c2: the precisions timed (the headline one must be among them)")
    ap.add_argument("--settle-max", type=float, default=4.0,
                    help="seconds of settling steps at most before a timed window (0: none, for profiler runs)")
    ap.add_argument("--ins-elem", choices=["tet4", "hexa8"], default="tet4", help="ins: element type")
    ap.add_argument("--rates", default="50,100,200,500,1000,2000")
    ap.add_argument("--serve-kinds", choices=["both", "vms", "supg"], default="both",
                    help="serve: request kinds (both: half VMS tet4, half SUPG hexa8)")
    ap.add_argument("--requests", type=int, default=120, help="requests per rate (serve)")
    ap.add_argument("--budget-gb", type=int, default=64)
    ap.add_argument("--max-active", type=int, default=64)
    ap.add_argument("--prep-threads", type=int, default=8)
    ap.add_argument("--chunk", type=int, default=25)
    ap.add_argument("--gpu-threads", type=int, default=4, help="serve: host threads issuing the resident requests' work")
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="serve: hardware queues of the process (HIP reads GPU_MAX_HW_QUEUES at start-up)")
    args = ap.parse_args()
    global SETTLE_MAX_S
    SETTLE_MAX_S = args.settle_max
    if args.steps is None:
        args.steps = {"c3": 500, "c1": 2000, "c2": 300, "c4": 200, "union": 200, "accuracy": 1000}.get(args.workload, 50)
    if args.workload == "serve" and args.hw_queues:
        # before the first HIP call: resident requests each run on their own stream
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(args.hw_queues, 32))
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only "
                                   "(--gpus must be 1)"}))
        sys.exit(2)
    if femx.device_count() < 1:
        print(json.dumps({"error": "no HIP device visible"}))
        sys.exit(1)
    if args.workload == "c4":
        workload_c4(args)
        return
    if args.workload == "serve":
        workload_serve(args)
        return
    if args.workload == "union":
        workload_union(args)
        return
    if args.workload == "c1":
        workload_c1(args)
        return
    if args.workload == "ins":
        workload_ins(args)
        return
    if args.workload == "c2":
        workload_c2(args)
        return
    if args.workload == "accuracy":
        workload_accuracy(args)
        return
    if args.workload == "hbm":
        workload_hbm(args)
        return
    if args.workload == "dropin":
        workload_dropin(args)
        return

    m, bc, params, ic = make_workload(args.n)
    E, N = m.n_elems, m.n_nodes
    r = N / E
    dm = femx.DeviceMesh(m)
    results = {}
    wanted = args.precisions.split(",")
    for prec_name, prec in (("mixed", femx.FEMX_MIXED), ("fp64", femx.FEMX_FP64),
                            ("mixed_f32sum", femx.FEMX_MIXED_F32SUM)):
        if prec_name not in wanted:
            continue
        secs, kt, finite, settled = time_solver(dm, bc, params, ic, prec, args.steps, args.warmup)
        results[prec_name] = dict(secs=secs, kt=kt, finite=finite, settled=settled)
    main_res = results[args.precision]
    mixed = args.precision == "mixed"
    ms_per_step = main_res["secs"] / args.steps * 1e3
    value = E * args.steps / main_res["secs"]

    # dominant kernel and its roofline
    kt = main_res["kt"]
    dom_name = max(kt, key=lambda k: kt[k]["total_ms"])
    dom = kt[dom_name]
    dom_ms = dom["total_ms"] / max(dom["launches"], 1)
    dom_bytes = dom["bytes_per_launch"]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    step_bytes = algorithmic_bytes_per_element_step(r, mixed) * E
    event_steps = max(args.steps, 200)
    step_ms_events = sum(v["total_ms"] for v in kt.values()) / event_steps
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(args.precision, args.n),
            "kernel": dom_name,
            "kernel_ms": round(dom_ms, 4), "kernel_bytes": dom_bytes,
            "step_achieved": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1),
            "step_frac": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "step_bytes_per_element": round(algorithmic_bytes_per_element_step(r, mixed), 2),
            "kernels": {k: dict(launches=v["launches"], mean_ms=round(v["total_ms"] / max(v["launches"], 1), 4),
                                bytes=v["bytes_per_launch"]) for k, v in kt.items()}}
    other = "fp64" if mixed else "mixed"
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "element-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-element/f64-state" if mixed else "f64",
        "data": "synthetic (seeded jittered Kuhn-split cube, channel BCs)",
        "config": {"workload": f"VMS tet4 {args.n}^3x6 channel Re200 jitter0.15 ({E} tets, {N} nodes)",
                   "elements": E, "nodes": N, "dual_iterations_per_step": 1,
                   "precision": args.precision},
        "roofline": roof,
        "finite_state": all(v["finite"] for v in results.values()),
        "event_step_ms": round(step_ms_events, 4),
        "settle": {"steps": main_res["settled"][0], "seconds": round(main_res["settled"][1], 3),
                   "rule": "after the warmup steps, 50-step chunks until two consecutive chunks agree within 1% "
                           f"(at least 0.5 s, at most {SETTLE_MAX_S:g} s); the kernel events come from a second pass of "
                           f"{event_steps} steps"},
    }
    for name in (other, "mixed_f32sum"):
        if name in results:
            out[name] = {"value": round(E * args.steps / results[name]["secs"], 1),
                         "ms_per_step": round(results[name]["secs"] / args.steps * 1e3, 4)}
    if "mixed_f32sum" in out:
        out["mixed_f32sum"]["note"] = F32SUM_NOTE
    ceil_gbs = stream_ceiling_gbs()
    roof["stream_ceiling"] = {"gbs": round(ceil_gbs, 1), "probe": CEILING_PROBE,
                              "achieved_frac": round(achieved / ceil_gbs, 4)}
    if not args.no_cpu:
        cv, csteps, csecs = cpu_baseline(m, bc, params, ic, args.cpu_budget)
        out["cpu_baseline"] = {"value": round(cv, 1), "unit": "element-steps/s", "cores": 1, "kind": "port",
                               "sample": f"{csteps} full FP64 steps of the same {E}-tet mesh, oracle "
                                         f"(plain-C restatement, 1 thread), {csecs:.1f} s"}
        out["speedup_vs_cpu_1t"] = round(value / cv, 1)
        nt = host_threads()
        if nt > 1:
            av, asteps, asecs = cpu_baseline(m, bc, params, ic, args.cpu_budget / 2, threads=nt)
            out["cpu_baseline_all_cores"] = {
                "value": round(av, 1), "unit": "element-steps/s", "cores": nt, "kind": "port",
                "sample": f"{asteps} full FP64 steps of the same mesh, oracle with its element sweeps on {nt} "
                          f"threads and the element-order scatter on one (bitwise the 1-thread results), "
                          f"{asecs:.1f} s"}
            out["speedup_vs_cpu_all_cores"] = round(value / av, 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
