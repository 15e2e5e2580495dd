"""Benchmark of the RK3 time step (compute_dt + compute_time_step_RK3,
main.f90:70-73) on one MI355X.

Workload: DNS at Re_b = 10000, 384x256x384 unique cells (faces
386x257x386), tanh y-stretching 2.2, CFL 0.5, seeded Poiseuille + 10%
perturbation projected divergence-free.  One "step" = one full RK3 step
(3 substeps with projections, boundary conditions, mass-flow correction and
the CFL time step).  Inputs are resident in HBM before the timed region.

Prints one JSON line: value = FP64 Mcell-updates/s; the mixed path
(FP32 stencil/FFT storage, FP64 Thomas/reductions) is reported beside it.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "channel-flow_amd"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def algorithmic_bytes(name, n_cells, esz, mh_ratio):
    """Unique input+output bytes per launch of each step kernel at storage
    precision (halo re-reads excluded).  mh_ratio = (nxp/2+1)*2/nxp: the
    half spectrum holds mh complex values per nxp reals."""
    spec = mh_ratio * n_cells * esz  # one pass over the half spectrum
    f = n_cells * esz                # one pass over a cell field
    table = {
        "rhs_rk": 9 * f,                 # read U,V,W (+G), write U',V',W' (+G)
        "div_fftx": 3 * f + spec,        # read U,V,W, write spectrum
        "fftz_fwd": 2 * spec,
        "fftz_inv": 2 * spec,
        "thomas": 2 * spec,
        "zsolve": 2 * spec,              # fused z transforms + y solve, both launches
        "zsolve_fwd": 2 * spec,          # its forward launch: read spectrum, write eliminated values
        "zsolve_bwd": 2 * spec,          # its backward launch: read them, write the solved spectrum
        "ifftx_proju": spec + f + 2 * f,  # read spectrum, write phi, read+write U
        "project_vw": 5 * f,             # read phi,V,W, write V,W
    }
    return table.get(name)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "rk3_step_384x256x384_fp64_hbm_traffic.md")


def measured_traffic(kernel, cells):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (rocprofv3 FETCH_SIZE / WRITE_SIZE runs scaled by a plain copy of known
    size); None when no summary covers this kernel and workload."""
    if cells != (384, 256, 384) or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        f = line.split()
        if len(f) == 4 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel and f[2] == "fp64":
            return int(f[3])
    return None


def run_precision(cfg, steps, warmup, cases):
    from channel_flow.solver import ChannelFlow
    g = ChannelFlow(cfg, device=0)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    del U, V, W
    log("%s: initialised, %.2f GB on device" % (cfg.precision, g.device_bytes() / 1e9))
    g.run(warmup, graph=True)
    g.sync()
    t0 = time.perf_counter()
    g.run(steps, graph=True)
    g.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    log("%s: %.3f ms/step (graph)" % (cfg.precision, ms))
    # per-kernel durations from HIP events on the solver's stream, same steps
    g.timing(True)
    g.timing_reset()
    t1 = time.perf_counter()
    g.run(steps, graph=False)
    g.sync()
    ms_ev = (time.perf_counter() - t1) * 1e3 / steps
    rep = g.timing_report()
    g.timing(False)
    _, div = g.check_divergence()
    return g, ms, ms_ev, rep, div


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """The cores this process may run on (its affinity mask, which is the
    host's CPU share on a GPU box), unless CF_CPU_THREADS says otherwise."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return int(os.environ.get("CF_CPU_THREADS", n))


def _oracle_step_time(cells, nth, max_steps, cap_s):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    from channel_flow import cases
    cfg = cases.case_dns384()
    cfg.cells = cells
    o = Oracle(cfg, nthreads=nth)
    U, V, W = cases.initial_field(cfg)
    o.initialise(U, V, W)
    del U, V, W
    n = 0
    t0 = time.perf_counter()
    while True:
        o.compute_dt()
        o.compute_time_step_RK3()
        n += 1
        el = time.perf_counter() - t0
        log("cpu baseline %s, %d threads: step %d done, %.1f s" % (cells, nth, n, el))
        if el > cap_s / 2 or n >= max_steps:
            break
    o.close()
    return el / n, n


def cpu_baseline(cells, steps_cap_s=30.0):
    """The oracle (FP64 C restatement of the reference step) on the host on a
    bounded sample of the workload: RK3 steps of the full grid with OpenMP
    over every core this process may use, and one step single-threaded on
    the same x-y planes with 1/8 of the z extent (the per-cell cost of the
    step does not depend on that extent, and a full-size single-threaded
    step would take minutes)."""
    nth = cpu_threads()
    sps, n = _oracle_step_time(cells, nth, 2, steps_cap_s)
    ncell = cells[0] * cells[1] * cells[2]
    small = (cells[0], cells[1], max(cells[2] // 8, 8))
    sps1, _ = _oracle_step_time(small, 1, 1, steps_cap_s)
    ncell1 = small[0] * small[1] * small[2]
    return {"value": round(ncell / sps / 1e6, 3), "unit": "Mcell-updates/s", "cores": nth, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": "%d RK3 step(s) of the %dx%dx%d workload on %d threads, %.2f s/step" % (n, *cells, nth, sps),
            "single_thread": {"value": round(ncell1 / sps1 / 1e6, 3), "unit": "Mcell-updates/s", "cores": 1,
                              "sample": "1 RK3 step of %dx%dx%d (the workload's x-y planes, 1/8 of its z "
                                        "extent), %.2f s/step" % (*small, sps1)}}


def other_configs(steps=50, warmup=5):
    """Step time of the other workloads (graph-replayed, inputs resident):
    64^3 DNS (FP64, fixed dt), 128x64x128 wall-modelled LES (Smagorinsky +
    EQWM, mixed and FP64), 256x96x256 wall-modelled LES at Re_tau 5200
    (mixed, EQWM), the reference's example input (32x40x32 cells, dynamic
    Smagorinsky, FP64) and the 128x64x128 LES with the dynamic model (mixed)."""
    from channel_flow import cases
    from channel_flow.solver import ChannelFlow
    res = {}
    dsm128 = cases.case_wmles128()
    dsm128.les_model = 2
    todo = [("dns64_fp64", cases.case_dns64(), "fp64"), ("wmles128_mixed", cases.case_wmles128(), "mixed"),
            ("wmles128_fp64", cases.case_wmles128(), "fp64"), ("wmles256_mixed", cases.case_wmles256(), "mixed"),
            ("example_dsm_fp64", cases.case_example(), "fp64"), ("wmles128_dsm_mixed", dsm128, "mixed")]
    for name, cfg, prec in todo:
        cfg.precision = prec
        cfg.wall_surrogate = 0
        g = ChannelFlow(cfg, device=0)
        g.initialise(*cases.initial_field(cfg))
        g.run(warmup, graph=True)
        g.sync()
        t0 = time.perf_counter()
        g.run(steps, graph=True)
        g.sync()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        n = cfg.cells[0] * cfg.cells[1] * cfg.cells[2]
        res[name] = {"cells": list(cfg.cells), "ms_per_step": round(ms, 4), "mcell_per_s": round(n / ms / 1e3, 1)}
        log("%s: %.3f ms/step" % (name, ms))
        g.close()
    return res


def batched_steps(members=8, steps=50, warmup=5):
    """The batched step (cf_batch_*) of one size class against the same cases
    stepped one after another: `members` DNS cases (FP64, Re_tau spread) on
    each of the serving ensemble's grids 64^3 and 96x64x96 and `members`
    wall-modelled LES cases (mixed) on 64^3, graph-replayed, inputs resident."""
    from channel_flow import cases
    from channel_flow.serving import ensemble_case
    from channel_flow.solver import ChannelFlow, ChannelFlowBatch
    res = {}
    for name, grid, wm in (("dns64", (64, 64, 64), False), ("dns96x64x96", (96, 64, 96), False),
                           ("wmles64", (64, 64, 64), True)):
        cfgs = [ensemble_case(grid, 180.0 + 250.0 * q, wm, name="b%d" % q) for q in range(members)]
        for q, c in enumerate(cfgs):
            c.seed = 40 + q
        b = ChannelFlowBatch(cfgs)
        for m, c in zip(b.members, cfgs):
            m.initialise(*cases.initial_field(c))
        b.run(warmup)
        t0 = time.perf_counter()
        b.run(steps)
        ms_b = (time.perf_counter() - t0) * 1e3 / steps
        b.close()
        singles = []
        for c in cfgs:
            g = ChannelFlow(c, device=0)
            g.initialise(*cases.initial_field(c))
            g.run(warmup, graph=True)
            singles.append(g)
        for g in singles:
            g.sync()
        t0 = time.perf_counter()
        for g in singles:
            g.run(steps, graph=True)
            g.sync()
        ms_s = (time.perf_counter() - t0) * 1e3 / steps
        for g in singles:
            g.close()
        n = grid[0] * grid[1] * grid[2] * members
        res[name] = {"members": members, "cells": list(grid), "batch_ms_per_step": round(ms_b, 4),
                     "one_after_another_ms_per_step": round(ms_s, 4),
                     "batch_mcell_per_s": round(n / ms_b / 1e3, 1), "gain": round(ms_s / ms_b, 3)}
        log("batched %s x%d: %.3f ms per batch step, %.3f ms one after another" % (name, members, ms_b, ms_s))
    return res


def serving_bench(scale, legs, budget_gb, chunk=25, n_cases=64, seed=5):
    """Ensemble of channel cases (seeded Poisson arrivals; grids 64^3,
    96x64x96, 128x96x128, 192x128x192; Re_tau ~ U[180, 2000]; DNS/WMLES
    50/50; 200-2000 steps times `scale`) through the serving loop: the same
    cases one at a time (a server with a single slot, same arrivals), then
    each leg of `legs`: ("streams", S) continuous batching over S per-case
    streams (cases running at once, one captured step graph each), or
    ("groups", S, B, C) cases of one size class stepped together by the
    batched step, groups of up to B members on S streams, stretches of at
    most `chunk` steps, only cases of at most C unique cells batched (0: all;
    larger cases on their own streams)."""
    from channel_flow import serving
    # steps and arrival gaps scale together: the offered load is the full ensemble's
    wl = serving.ensemble_workload(n=n_cases, seed=seed, steps_scale=scale, mean_gap=0.05 * scale)
    fields = serving._FieldCache()
    for cfg, _, _ in wl:
        fields(cfg)
    log("serving: %d cases, %d steps in total, fields ready" % (len(wl), sum(n for _, n, _ in wl)))
    budget = int(budget_gb * 1e9)
    t0 = time.perf_counter()
    recs1, summ1 = serving.serve(wl, budget, max_active=1, fields=fields)
    seq = serving.latency_report(recs1, summ1)
    log("serving: one-at-a-time run done in %.1f s: %s" % (time.perf_counter() - t0, seq))
    if seq["done"] != seq["cases"]:
        raise RuntimeError("one-at-a-time serving run finished %d of %d cases" % (seq["done"], seq["cases"]))
    runs = {}
    for leg in legs:
        t0 = time.perf_counter()
        if leg[0] == "streams":
            key = "streams_%d" % leg[1]
            recs, summ = serving.serve(wl, budget, max_active=leg[1], fields=fields)
        else:
            _, sl, b, mc = leg
            key = "groups_%dx%d" % (sl, b) + ("_upto_%d_cells" % mc if mc else "")
            recs, summ = serving.serve(wl, budget, max_active=sl, fields=fields, batch=b, chunk=chunk,
                                       batch_max_cells=mc)
        runs[key] = serving.latency_report(recs, summ)
        log("serving: %s done in %.1f s: %s" % (key, time.perf_counter() - t0, runs[key]))
        if runs[key]["done"] != runs[key]["cases"]:
            raise RuntimeError("serving run %s finished %d of %d cases" % (key, runs[key]["done"], runs[key]["cases"]))
    best = max(runs, key=lambda k: runs[k]["mcell_per_s"])
    mix = runs[best]
    return {"workload": "%d cases, Poisson arrivals mean gap %g ms, steps %d-%d per case" % (
                n_cases, 50 * scale, max(1, round(200 * scale)), max(1, round(2000 * scale))),
            "steps_scale": scale,
            "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "default (4)"),
            "best": best, "hbm_budget_gb": budget_gb, "group_chunk": chunk,
            "batched": mix, "one_at_a_time": seq,
            "runs": runs,
            "throughput_gain": round(mix["mcell_per_s"] / seq["mcell_per_s"], 3),
            "case_p99_gain": round(seq["case_s_p99"] / mix["case_s_p99"], 3)}


def parse_legs(spec):
    """'streams:8,groups:8:4:0' -> [("streams", 8), ("groups", 8, 4, 0)]"""
    out = []
    for item in spec.split(","):
        f = item.split(":")
        if f[0] == "streams" and len(f) == 2:
            out.append(("streams", int(f[1])))
        elif f[0] == "groups" and len(f) in (3, 4):
            out.append(("groups", int(f[1]), int(f[2]), int(f[3]) if len(f) == 4 else 0))
        else:
            raise ValueError("bad serving leg %r" % item)
    return out


def serving_child(a):
    """The serving leg in a child process, whose HIP runtime can open
    --serve-hw-queues hardware queues (GPU_MAX_HW_QUEUES is read when the
    runtime starts)."""
    import subprocess
    env = dict(os.environ)
    if a.serve_hw_queues > 0:
        env["GPU_MAX_HW_QUEUES"] = str(min(a.serve_hw_queues, 32))
    cmd = [sys.executable, os.path.abspath(__file__), "--serve-only", "--serve-scale", str(a.serve_scale),
           "--serve-legs", a.serve_legs, "--serve-budget-gb", str(a.serve_budget_gb),
           "--serve-chunk", str(a.serve_chunk)]
    r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, text=True, timeout=1200)
    if r.returncode != 0:
        raise RuntimeError("serving leg failed with status %d" % r.returncode)
    return json.loads(r.stdout.strip().splitlines()[-1])


def copy_calibration(mib):
    """A plain device-to-device copy of known size (torch's copy kernel),
    for scaling the HBM byte counters (FETCH_SIZE, WRITE_SIZE) of a profiler
    run: each launch reads and writes exactly `mib` MiB."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "launches": 5}))


def measured_copy_gbps(mib=1024, reps=10):
    """Read+write rate of a plain 1 GiB device copy on this GPU (torch's copy,
    HIP events on torch's stream): the practical ceiling a streaming kernel
    that reads and writes is compared with, beside the spec peak."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2.0 * n * 8 / (ms / 1e3) / 1e9


def beside_step_timings(g, cells, reps=3):
    """The rows beside the step, on the FP64 context of the main run: the
    actual pressure (compute_pressure, pressure.f90:21-341) and one plane
    statistics accumulation (statistics.f90:22-171) from HIP events on the
    solver's stream (device time, excluding the host copy of P), and the
    snapshot writer's host rate (input_output.f90:393-544 layout)."""
    import tempfile
    from channel_flow.snapshot import write_snapshot
    g.timing(True)
    g.timing_reset()
    t0 = time.perf_counter()
    for _ in range(reps):
        P = g.compute_pressure()
    t_p = (time.perf_counter() - t0) / reps
    for _ in range(reps):
        g.stats_accumulate()
    g.sync()
    rep = g.timing_report()
    g.timing(False)
    U, V, W = g.get_state()
    nbytes = 8 * (U.size + V.size + W.size + P.size)
    with tempfile.TemporaryDirectory() as td:
        t0 = time.perf_counter()
        write_snapshot(os.path.join(td, "snap"), g.x, g.y, g.z, U, V, W, P=P)
        t_w = time.perf_counter() - t0
    n = cells[0] * cells[1] * cells[2]
    cp = rep.get("compute_pressure", (None, 0))[0]
    sa = rep.get("stats_accumulate", (None, 0))[0]
    return {"compute_pressure_device_ms": round(cp, 4) if cp else None,
            "compute_pressure_call_s": round(t_p, 3),
            "compute_pressure_mcell_per_s": round(n / cp / 1e3, 1) if cp else None,
            "stats_accumulate_device_ms": round(sa, 4) if sa else None,
            "snapshot_write_GBps": round(nbytes / t_w / 1e9, 2)}


def dsm_steps(cells, steps, warmup=2):
    """The dynamic Smagorinsky model (subgrid.f90:376-538) on the DNS
    workload's grid: RK3 steps with les_model 2 (one model evaluation per
    substep), graph-replayed, and the model's share from HIP events."""
    from channel_flow import cases
    from channel_flow.solver import ChannelFlow
    cfg = cases.case_dns384("fp64")
    cfg.cells = cells
    cfg.les_model = 2
    g = ChannelFlow(cfg, device=0)
    g.initialise(*cases.initial_field(cfg))
    g.run(warmup, graph=True)
    g.sync()
    t0 = time.perf_counter()
    g.run(steps, graph=True)
    g.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    g.timing(True)
    g.timing_reset()
    g.run(steps, graph=False)
    g.sync()
    rep = g.timing_report()
    g.close()
    n = cells[0] * cells[1] * cells[2]
    ev = rep.get("dynamic_smagorinsky", (None, 0))
    log("dsm %s: %.3f ms/step, %.3f ms per model evaluation" % (cells, ms, ev[0] or float("nan")))
    # algorithmic bytes of one evaluation: read U, V, W; write nu_t
    ab = 4 * n * 8
    return {"cells": list(cells), "ms_per_step": round(ms, 4), "mcell_per_s": round(n / ms / 1e3, 1),
            "model_ms_per_eval": round(ev[0], 4) if ev[0] else None, "evals": ev[1],
            "model_GBps_on_alg_bytes": round(ab / (ev[0] / 1e3) / 1e9, 1) if ev[0] else None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=str, default="384,256,384")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-others", action="store_true", help="skip the 64^3 and LES step timings")
    ap.add_argument("--serve-scale", type=float, default=1.0,
                    help="steps of the serving ensemble relative to 200-2000 per case")
    ap.add_argument("--serve-legs", type=str, default="streams:8,groups:8:4:0,groups:8:4:1572864",
                    help="serving runs beside the one-at-a-time run: streams:S (S per-case streams) or "
                         "groups:S:B[:C] (size-class groups of up to B members on S streams, cases of at most "
                         "C unique cells batched, 0 all)")
    ap.add_argument("--serve-chunk", type=int, default=25, help="steps per stretch of a group")
    ap.add_argument("--serve-hw-queues", type=int, default=0,
                    help="hardware queues of the serving process (GPU_MAX_HW_QUEUES, at most 32; 0: the "
                         "runtime's default).  Measured at 64 cases: 16 queues with 4, 8 or 16 slots gave "
                         "4411, 4597 and 4669 Mcell/s against 4616 and 4721 with the default 4 queues and "
                         "4 or 8 slots: the GPU, not the queues, bounds the mix")
    ap.add_argument("--serve-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain copy of MIB MiB five times (counter calibration)")
    ap.add_argument("--copy-ceiling", action="store_true", help="only measure a 1 GiB device copy")
    ap.add_argument("--batch-only", action="store_true", help="only time the batched step of one size class")
    ap.add_argument("--dsm-only", action="store_true",
                    help="only time RK3 steps with the dynamic Smagorinsky model on --cells")
    a = ap.parse_args()
    if a.copy_ceiling:
        print(json.dumps({"copy_GBps": measured_copy_gbps()}))
        return
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.gpus != 1:
        print("bench.py measures one GPU: the solver runs a single case per GPU and has no "
              "multi-GPU path; use --gpus 1", file=sys.stderr)
        sys.exit(2)
    from channel_flow import cases
    cells = tuple(int(c) for c in a.cells.split(","))
    if a.serve_only:
        print(json.dumps(serving_bench(a.serve_scale, parse_legs(a.serve_legs), a.serve_budget_gb, a.serve_chunk)))
        return
    if a.dsm_only:
        print(json.dumps({"dsm": dsm_steps(cells, a.steps, a.warmup)}))
        return
    if a.batch_only:
        print(json.dumps({"batched_steps": batched_steps()}))
        return
    ncell = cells[0] * cells[1] * cells[2]
    mh_ratio = (cells[0] // 2 + 1) * 2.0 / cells[0]

    cfg = cases.case_dns384("fp64")
    cfg.cells = cells
    g64, ms64, ms64_ev, rep64, div64 = run_precision(cfg, a.steps, a.warmup, cases)
    value = ncell / (ms64 / 1e3) / 1e6
    # dominant kernel and its roofline
    tot = {k: v[0] * v[1] for k, v in rep64.items()}
    dom = max(tot, key=tot.get)
    dom_ms = rep64[dom][0]
    ab = algorithmic_bytes(dom, ncell, 8, mh_ratio)
    achieved = ab / (dom_ms / 1e3) / 1e9 if ab else None
    kernels = {k: {"avg_ms": round(v[0], 4), "launches": v[1],
                   "GBps": (round(algorithmic_bytes(k, ncell, 8, mh_ratio) / (v[0] / 1e3) / 1e9, 1)
                            if algorithmic_bytes(k, ncell, 8, mh_ratio) else None)}
               for k, v in rep64.items()}
    out = {
        "metric": "Mcell-updates/s per RK3 step, FP64 vs mixed, 384x256x384; ensemble p50/p99 latency",
        "value": round(value, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms64, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded Poiseuille + 10% perturbation, projected)",
        "config": {"workload": "DNS Re_b=10000 %dx%dx%d tanh2.2 CFL0.5" % cells,
                   "cells": list(cells), "precision": "fp64", "graph": True},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1) if achieved else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "algorithmic_bytes": ab, "traffic": measured_traffic(dom, cells)},
        "copy_ceiling": None,
        "kernels_fp64": kernels,
        "ms_per_step_event_timed": round(ms64_ev, 4),
        "max_div_fp64": div64,
    }
    # in a child process: torch brings its own HIP runtime, which must not
    # share a process with the one the library is linked against
    try:
        import subprocess
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-ceiling"], capture_output=True,
                           text=True, timeout=300)
        cg = json.loads(r.stdout.strip().splitlines()[-1])["copy_GBps"]
        out["copy_ceiling"] = {"GBps": round(cg, 1), "frac_of_peak": round(cg / HBM_PEAK_GBS, 4),
                               "dominant_kernel_frac_of_copy": round(achieved / cg, 4) if achieved else None,
                               "what": "1 GiB device copy, read+write bytes / HIP-event time"}
    except Exception as e:
        log("copy ceiling not measured: %s" % e)
    if not a.no_others:
        out["beside_step_fp64"] = beside_step_timings(g64, cells)
    if not a.no_mixed:
        U64, V64, W64 = g64.get_state()
        s64 = g64.compute_statistics()
        g64.close()
        cfgm = cases.case_dns384("mixed")
        cfgm.cells = cells
        gm, msm, msm_ev, repm, divm = run_precision(cfgm, a.steps, a.warmup, cases)
        Um, Vm, Wm = gm.get_state()
        sm = gm.compute_statistics()
        err = max(float(np.max(np.abs(x - y)) / np.max(np.abs(y))) for x, y in
                  ((Um, U64), (Vm, V64), (Wm, W64)))
        prof = float(np.max(np.abs(sm["Umean"] - s64["Umean"])) / np.max(np.abs(s64["Umean"])))
        out["mixed"] = {"value": round(ncell / (msm / 1e3) / 1e6, 2), "ms_per_step": round(msm, 4),
                        "speedup_vs_fp64": round(ms64 / msm, 3), "max_div": divm,
                        "field_max_rel_err_vs_fp64": err, "mean_profile_rel_err": prof,
                        "steps_compared": a.steps * 2 + a.warmup,
                        "kernels": {k: {"avg_ms": round(v[0], 4), "launches": v[1]} for k, v in repm.items()}}
        gm.close()
    else:
        g64.close()
    if not a.no_serve:
        out["serving"] = serving_child(a)
    if not a.no_others:
        out["other_configs"] = other_configs()
        out["other_configs"]["dns384_dsm_fp64"] = dsm_steps(cells, 5)
        out["batched_steps"] = batched_steps()
    if not a.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(cells)
        except MemoryError:
            out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
