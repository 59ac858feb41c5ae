"""Throughput of the fused per-level advance on one MI355X.

Workload: the 4096x4096 FP64 plain-Roe dam break (flat bed, no friction,
CFL 0.45, transmissive sides).  One step = the fused advance kernel, which
also reduces the CFL estimate of the next step on the device (no host sync).
Inputs are resident in HBM before the timed region starts.

Prints one JSON line: cell-updates/s, the HBM-roofline fraction of the
advance kernel (algorithmic 64 B per cell-update over its average duration,
HIP events on the kernel's stream) and the oracle's CPU throughput on a
bounded sample of the same workload.

The headline uses the fast arithmetic build (shared reciprocals, FMA
contraction); the exact build (bit-identical to the FP64 oracle) is timed on
the same workload, and the relative L-infinity distance between the two
final states after all timed steps is reported beside the numbers.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CONDITION_MS = 60.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096, help="grid side of the dam break")
    ap.add_argument("--dtype", choices=["f64", "f32"], default="f64")
    ap.add_argument("--solver", choices=["Roe", "HydrostaticRoe"], default="Roe")
    ap.add_argument("--rows", type=int, default=0, help="strip rows per wavefront (0 = default)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="budget of the five timed CPU-baseline samples, per thread count")
    ap.add_argument("--condition-ms", type=float, default=CONDITION_MS,
                    help="HBM streaming right before each timed region's warm-up (0: none)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", dest="extra", action="store_false",
                    help="skip the other advance workloads (bumpy bed, FP32, lakes at rest)")
    ap.add_argument("--cpu-worker", nargs=2, default=None, help=argparse.SUPPRESS)
    ap.add_argument("--arith", choices=["fast", "exact"], default="fast")
    ap.add_argument("--only", choices=["ensemble", "adjoint", "adjoint_ensemble", "two_level", "amr"],
                    default=None,
                    help="run one secondary leg alone and print its JSON (for profiling)")
    ap.add_argument("--adjoint-steps", type=int, default=400,
                    help="steps of the 2048^2 calibration rollout timed forward + adjoint (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=120,
                    help="timed steps of the batched ensemble and of the per-scenario loop (0: skip)")
    ap.add_argument("--adjoint-ensemble-steps", type=int, default=100,
                    help="steps of the batched forward + adjoint leg (0: skip)")
    ap.add_argument("--two-level-steps", type=int, default=100,
                    help="timed coarse steps of the two-level dam break (0: skip)")
    ap.add_argument("--amr-steps", type=int, default=100,
                    help="timed coarse steps of the 3-level cylinder dam break hierarchy (0: skip)")
    return ap.parse_args()


def _cpu_worker(n: int, budget_s: float) -> None:
    """Child process of cpu_baseline (no torch, no HIP): time the oracle on
    the n x n dam break and print one JSON line.  One warm-up step (which
    also sizes the samples), then 5 samples of k whole steps each, k chosen
    so the samples take about budget_s together; the median sample counts."""
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "hx_workloads", os.path.join(ROOT, "hydroexa_amd", "workloads.py"))
    W = importlib.util.module_from_spec(spec)
    sys.modules["hx_workloads"] = W          # dataclasses look their module up
    spec.loader.exec_module(W)
    import oracle as O
    w = W.dam_break(n, 1)
    args = (w.z, w.n, w.dx, w.dy, w.cfl)
    t0 = time.perf_counter()
    O.run(O.ROE, w.h, w.hu, w.hv, *args, 1, w.bc_lo, w.bc_hi)
    per_step = time.perf_counter() - t0
    k = int(max(1, min(200, round(budget_s / 5 / max(per_step, 1e-9)))))
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        O.run(O.ROE, w.h, w.hu, w.hv, *args, k, w.bc_lo, w.bc_hi)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    print(json.dumps({"value": w.cells * k / med / 1e6, "steps_per_sample": k,
                      "sample_s": times, "name": w.name, "cells": w.cells}), flush=True)


def _cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            return next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except (OSError, StopIteration):
        return "unknown"


def cpu_baseline(n, budget_s):
    """The oracle (FP64 C/OpenMP restatement) on the host cores, in a child
    process so that the OpenMP placement takes effect: OMP_NUM_THREADS
    threads (the host's affinity when unset), OMP_PROC_BIND=close,
    OMP_PLACES=cores; then the same on one thread.  Median of 5 samples after
    a warm-up step (the baseline protocol of the measurement plan)."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))

    def run(nthreads, budget):
        env = dict(os.environ, OMP_NUM_THREADS=str(nthreads), OMP_PROC_BIND="close",
                   OMP_PLACES="cores")
        res = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-worker",
                              str(n), str(budget)], env=env, capture_output=True, text=True,
                             timeout=600)
        if res.returncode != 0:
            raise RuntimeError("CPU baseline worker failed: " + res.stderr[-2000:])
        return json.loads(res.stdout.strip().splitlines()[-1])

    many = run(threads, budget_s)
    one = run(1, budget_s)
    k, k1 = many["steps_per_sample"], one["steps_per_sample"]
    return {"value": many["value"], "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"median of 5 samples of {k} steps of the {many['name']} workload "
                      f"({many['cells']} cells) after a 1-step warm-up, oracle with {threads} "
                      f"OpenMP threads (OMP_PROC_BIND=close, OMP_PLACES=cores); samples "
                      + ", ".join(f"{t:.2f}" for t in many["sample_s"]) + " s",
            "one_core": {"value": one["value"], "cores": 1,
                         "sample": f"median of 5 samples of {k1} steps on one thread; samples "
                                   + ", ".join(f"{t:.2f}" for t in one["sample_s"]) + " s"},
            "cpu_model": _cpu_model()}


def condition_device(ms=CONDITION_MS, cells=1 << 22):
    """Keep the device streaming HBM for about `ms` ms (hx_stream_probe on
    its own buffers, no state of the workload touched).  After an idle
    period -- the host-to-device upload of a level -- the device runs an
    HBM-bound kernel at a clock state that differs from its sustained one for
    about 30 steps of the 4096^2 advance (the same on a lake at rest with no
    flow at all), so a short timed region that starts right after an upload
    measures that transient.  Running this just before the warm-up steps
    puts the timed steps at the sustained state a long simulation runs at.
    Returns the milliseconds spent."""
    import ctypes as C
    import torch
    from hydroexa_amd import _lib as L

    if ms <= 0:
        return 0.0
    lib = L.load()
    ctx = C.c_void_p()
    L.check(lib.hx_ctx_create(C.byref(ctx), torch.cuda.current_device()), "hx_ctx_create")
    try:
        planes = [torch.full((cells,), 0.5 + k, dtype=torch.float64, device="cuda") for k in range(8)]
        ins = (C.c_void_p * 5)(*[t.data_ptr() for t in planes[:5]])
        outs = (C.c_void_p * 3)(*[t.data_ptr() for t in planes[5:]])
        sp = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        while (time.perf_counter() - t0) * 1e3 < ms:
            for _ in range(50):
                L.check(lib.hx_stream_probe(ctx, cells, ins, outs, sp), "hx_stream_probe")
            torch.cuda.synchronize()
        spent = (time.perf_counter() - t0) * 1e3
        del planes
    finally:
        lib.hx_ctx_destroy(ctx)
    return spent


def time_gpu(w, dtype, steps, warmup, rows, arith="fast", keep=False, condition_ms=CONDITION_MS):
    import torch
    from hydroexa_amd import Level, SolverVariant, set_arith, set_tuning

    tdt = torch.float64 if dtype == "f64" else torch.float32
    set_tuning(w.solver, tdt, rows, 0)
    set_arith(arith)
    lv = Level(w.h, w.hu, w.hv, w.z, w.n, w.dx, w.dy, w.bc_lo, w.bc_hi, dtype=tdt)
    sv = SolverVariant(w.solver)
    condition_device(condition_ms)
    sv.run(lv, w.cfl, warmup)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(st)
    # the timed steps continue the warm-up rollout: the first dt comes from
    # the estimate the last warm-up step reduced (no standalone CFL pass)
    dth, vh = sv.run(lv, w.cfl, steps, resume=warmup > 0)
    e1.record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = e0.elapsed_time(e1)
    state = lv.U_new.clone() if keep else None
    del lv
    set_arith("exact")
    return ev_ms, wall, dth, state


def traffic_from_profile(name):
    """HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE (calibrated against a copy of known size), when a summary for
    this kernel and workload is committed under profiles/."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        d = json.load(open(path))
    except (OSError, ValueError):
        return None
    e = d.get(name)
    return e.get("bytes_per_launch") if isinstance(e, dict) else None


def stream_ceiling(cells, reps=20):
    """Streaming ceiling of the advance's access pattern: the library's
    elementwise probe that reads five FP64 planes of `cells` elements and
    writes three (hx_stream_probe), timed with HIP events on the stream it
    runs on.  Returns ms per pass and GB/s at 64 B per element."""
    import ctypes as C
    import torch
    from hydroexa_amd import _lib as L

    lib = L.load()
    ctx = C.c_void_p()
    L.check(lib.hx_ctx_create(C.byref(ctx), torch.cuda.current_device()), "hx_ctx_create")
    try:
        planes = [torch.full((cells,), 0.5 + k, dtype=torch.float64, device="cuda") for k in range(8)]
        ins = (C.c_void_p * 5)(*[t.data_ptr() for t in planes[:5]])
        outs = (C.c_void_p * 3)(*[t.data_ptr() for t in planes[5:]])
        st = torch.cuda.current_stream()
        sp = C.c_void_p(st.cuda_stream)
        for _ in range(3):
            L.check(lib.hx_stream_probe(ctx, cells, ins, outs, sp), "hx_stream_probe")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            lib.hx_stream_probe(ctx, cells, ins, outs, sp)
        e1.record(st)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        del planes
    finally:
        lib.hx_ctx_destroy(ctx)
    return {"kernel": "stream5r3w (read 5 FP64 planes, write 3, elementwise)", "cells": cells,
            "ms": ms, "GBps": 64.0 * cells / (ms / 1e3) / 1e9}


def time_conservation(w, dtype, reps=20):
    """Fused conservation reduction (mass, momenta, KE, min h) of the
    workload's level, wall time per call including the host sync of the
    5-value result; 4 words read per cell (h, hu, hv, roughness)."""
    import torch
    from hydroexa_amd import Level
    from hydroexa_amd.diagnostics import level_conservation

    dt = torch.float64 if dtype == "f64" else torch.float32
    lv = Level(w.h, w.hu, w.hv, w.z, w.n, w.dx, w.dy, w.bc_lo, w.bc_hi, dtype=dt)
    vals = level_conservation(lv)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        vals = level_conservation(lv)
    ms = (time.perf_counter() - t0) / reps * 1e3
    es = 8 if dtype == "f64" else 4
    return {"kernel": "conservation reduction", "cells": w.cells, "ms_per_call": ms,
            "GBps_incl_sync": 4 * es * w.cells / (ms / 1e3) / 1e9,
            "mass": float(vals[0]), "min_h": float(vals[4])}


def time_ensemble(steps, arith="fast"):
    """256 dam-break scenarios of uneven size (FP32): the batched ensemble
    (one launch per step for all scenarios, 6-step HIP graph replayed,
    autotuned launch shape) against the plain loop that runs each scenario
    on its own, from the same initial states and for the same steps."""
    import torch
    from hydroexa_amd import Ensemble, Level, SolverVariant, set_arith
    from hydroexa_amd.autotune import tune_ensemble
    from hydroexa_amd import workloads as W

    set_arith(arith)
    ws = W.ensemble(256)
    cells = sum(w.cells for w in ws)

    def make():
        return [Level(w.h, w.hu, w.hv, w.z, w.n, w.dx, w.dy, w.bc_lo, w.bc_hi,
                      dtype=torch.float32) for w in ws]

    levels = make()
    ens = Ensemble(levels, "Roe")
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        best, _ = tune_ensemble(ens, steps=6, rows=(8, 16, 32), occs=(1,), use_cache=False, stream=st)
        ens.run(0.45, 6, graph=True, stream=st)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        ens.run(0.45, steps, graph=True, stream=st)
        e1.record(st)
    torch.cuda.synchronize()
    batched_ms = e0.elapsed_time(e1)
    ens.close()
    del levels, ens
    torch.cuda.empty_cache()
    plain = make()
    sv = SolverVariant("Roe")
    for lv in plain[:8]:
        sv.run(lv, 0.45, 2)             # warm-up of the single-level path
    plain = make()
    torch.cuda.synchronize()
    cur = torch.cuda.current_stream()
    p0, p1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    p0.record(cur)
    for lv in plain:
        sv.run(lv, 0.45, steps)
    p1.record(cur)
    torch.cuda.synchronize()
    plain_ms = p0.elapsed_time(p1)
    set_arith("exact")
    del plain
    torch.cuda.empty_cache()
    return {"workload": "ensemble of 256 dam breaks, sides U{128..1024} (seeded), FP32, plain Roe",
            "cells": cells, "steps": steps, "arith": arith,
            "tuning": {"rows": best[0], "occupancy": best[1], "persistent": best[2]},
            "batched_Mcups": cells * steps / (batched_ms / 1e3) / 1e6,
            "plain_loop_Mcups": cells * steps / (plain_ms / 1e3) / 1e6,
            "speedup": plain_ms / batched_ms,
            "batched_ms_per_step": batched_ms / steps, "plain_ms_per_step": plain_ms / steps}


def time_adjoint(steps, every=20, arith="fast"):
    """Forward rollout and hand-written adjoint on the calibration workload:
    2048^2 FP64 hydrostatic Roe with Manning friction, 64 hills and a NODATA
    margin, 32 gauges, checkpoint every `every` steps, dt frozen from the
    first forward.  Times one forward (gauge series + checkpoints) and one
    backward (segment recomputation + adjoint steps) with HIP events, and
    one full calibration iteration through the autograd Function."""
    import torch
    from hydroexa_amd import set_arith
    from hydroexa_amd.pyhydroexa import synthetic_problem

    set_arith(arith)
    prob = synthetic_problem(n=2048, nsteps=steps, ckpt_every=every)
    ro = prob.rollout
    cells = prob.n * prob.n
    U0 = torch.stack([prob.h0, prob.hu0, prob.hv0]).contiguous()
    n_ctrl = torch.full((prob.ctrl, prob.ctrl), 0.04, dtype=torch.float64, device=U0.device)
    ter = torch.stack([prob.bed(), prob.roughness(n_ctrl)]).contiguous()
    st = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    series, _ = ro.forward_raw(U0, ter)                        # warm-up pass
    gs = 2.0 * (series - prob.observed) / series.numel()
    ro.backward_raw(gs, None)
    torch.cuda.synchronize()
    ev[0].record(st)
    series, _ = ro.forward_raw(U0, ter)
    ev[1].record(st)
    gs = 2.0 * (series - prob.observed) / series.numel()
    g_state, g_ter = ro.backward_raw(gs, None)
    ev[2].record(st)
    torch.cuda.synchronize()
    fwd_ms, bwd_ms = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
    # one calibration iteration through the autograd Function
    nc = n_ctrl.clone().requires_grad_(True)
    t0 = time.perf_counter()
    loss = prob.loss(nc)
    loss.backward()
    torch.cuda.synchronize()
    it_s = time.perf_counter() - t0
    ncp = (steps + every - 1) // every
    out = {"workload": "calibration rollout 2048x2048 FP64 hydrostatic Roe + Manning, 64 hills, "
                       "NODATA margin 32, 32 gauges, 64x64 roughness control grid",
           "cells": cells, "steps": steps, "ckpt_every": every, "arith": arith,
           "forward_ms": fwd_ms, "backward_ms": bwd_ms,
           "forward_Mcups": cells * steps / (fwd_ms / 1e3) / 1e6,
           "adjoint_Mcups": cells * steps / (bwd_ms / 1e3) / 1e6,
           "forward_plus_adjoint_Mcups": cells * steps / ((fwd_ms + bwd_ms) / 1e3) / 1e6,
           "backward_over_forward": bwd_ms / fwd_ms,
           "recomputed_steps": steps,   # hydrostatic: every segment's last step too
           "checkpoint_bytes": ncp * 3 * cells * 8,
           "calibration_iteration_s": it_s, "loss": float(loss.item()),
           "grad_n_ctrl_norm": float(nc.grad.norm().item())}
    ro.close()
    set_arith("exact")
    del prob, U0, ter, g_state, g_ter
    torch.cuda.empty_cache()
    return out


def time_adjoint_ensemble(steps, nscen=128, every=10, arith="fast"):
    """Batched forward + adjoint (hx_adjoint_batch_*) against the loop that
    runs each scenario's forward + adjoint on its own (hx_adjoint_*, the same
    kernels one scenario at a time): workloads.adjoint_ensemble, nscen dam
    breaks of sides U{128..512} (seeded), hydrostatic Roe + Manning, FP64,
    `steps` steps with a checkpoint every `every`, 8 gauges per scenario,
    loss = sum of squares of the gauge series.  Both take the same CFL dt
    and give the same bits (checked here on every scenario's roughness
    gradient).  Cell updates of forward + adjoint = cells x steps."""
    import torch
    from hydroexa_amd import set_arith
    from hydroexa_amd import workloads as W
    from hydroexa_amd.pyhydroexa import Rollout, RolloutBatch

    set_arith(arith)
    ws, gauges = W.adjoint_ensemble(nscen, 128, 512, steps=steps)
    cells = sum(w.cells for w in ws)
    T = lambda a: torch.as_tensor(a, dtype=torch.float64, device="cuda")
    U0s = [torch.stack([T(w.h), T(w.hu), T(w.hv)]).contiguous() for w in ws]
    Ts = [torch.stack([T(w.z), T(w.n)]).contiguous() for w in ws]
    sc = [dict(nx=w.h.shape[1], ny=w.h.shape[0], dx=w.dx, dy=w.dy, bc_lo=w.bc_lo, bc_hi=w.bc_hi,
               gauges=g) for w, g in zip(ws, gauges)]
    st = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    rb = RolloutBatch("HydrostaticRoe", sc, steps, ckpt_every=every)
    for rep in range(2):                      # the first pass warms up
        ev[0].record(st)
        series, _ = rb.forward_raw(U0s, Ts)
        ev[1].record(st)
        g_state, g_ter = rb.backward_raw([2.0 * x for x in series], None)
        ev[2].record(st)
        torch.cuda.synchronize()
    b_fwd, b_bwd = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
    ros = [Rollout("HydrostaticRoe", s["nx"], s["ny"], s["dx"], s["dy"], s["bc_lo"], s["bc_hi"],
                   steps, s["gauges"], ckpt_every=every) for s in sc]
    for k in range(min(4, nscen)):            # warm-up of the one-scenario path
        se, _ = ros[k].forward_raw(U0s[k], Ts[k])
        ros[k].backward_raw(2.0 * se, None)
    torch.cuda.synchronize()
    same = True
    ev[0].record(st)
    outs = []
    for k, ro in enumerate(ros):
        se, _ = ro.forward_raw(U0s[k], Ts[k])
        gs1, gt1 = ro.backward_raw(2.0 * se, None)
        outs.append(gt1)
    ev[1].record(st)
    torch.cuda.synchronize()
    loop_ms = ev[0].elapsed_time(ev[1])
    for k in range(nscen):
        same = same and torch.equal(outs[k], g_ter[k])
    for ro in ros:
        ro.close()
    rb.close()
    set_arith("exact")
    del ros, rb, outs, g_state, g_ter, U0s, Ts
    torch.cuda.empty_cache()
    batch_ms = b_fwd + b_bwd
    return {"workload": f"batched forward + adjoint: {nscen} dam breaks, sides U{{128..512}} "
                        f"(seeded), hydrostatic Roe + Manning, FP64, {steps} steps, checkpoint "
                        f"every {every}, 8 gauges each",
            "cells": cells, "steps": steps, "arith": arith,
            "batched_forward_ms": b_fwd, "batched_backward_ms": b_bwd,
            "batched_Mcups": cells * steps / (batch_ms / 1e3) / 1e6,
            "per_scenario_loop_ms": loop_ms,
            "per_scenario_loop_Mcups": cells * steps / (loop_ms / 1e3) / 1e6,
            "speedup": loop_ms / batch_ms, "gradients_bitwise_equal": bool(same)}


def time_two_level(steps, arith="fast"):
    """Two-level coupled step (ratio 2, subcycled): the 4096^2 FP64 dam break
    with the 1024^2 coarse cells around the dam refined (2048^2 fine cells),
    plain Roe.  One step = dt from both levels' CFL, coarse register +
    advance, two fine substeps (ghost fill, register, advance), reflux,
    average_down.  Timed as the device-resident run (hx_two_level_run) and
    as host-driven steps (CFL read back per level each step).  Composite
    cell updates per step = coarse cells + 2 x fine cells.  Also times the
    coupling kernels alone (fills, registers, reflux, average_down) with
    HIP events."""
    import ctypes as C
    import torch
    from hydroexa_amd import Level, TwoLevel, set_arith
    from hydroexa_amd import _lib as L
    from hydroexa_amd import workloads as W
    from hydroexa_amd.solver import context

    set_arith(arith)
    w = W.dam_break(4096, 1)
    lv = Level(w.h, w.hu, w.hv, w.z, w.n, w.dx, w.dy, w.bc_lo, w.bc_hi)
    box = (1536, 1536, 1024, 1024)
    T = TwoLevel("Roe", lv, box)
    T.run(3, w.cfl)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # device-resident run (hx_two_level_run)
    t0 = time.perf_counter()
    e0.record(st)
    T.run(steps, w.cfl)
    e1.record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / steps
    # host-driven steps (per-level CFL read back each step), for comparison
    hs = max(steps // 4, 1)
    e0.record(st)
    for _ in range(hs):
        T.step(w.cfl)
    e1.record(st)
    torch.cuda.synchronize()
    host_ms = e0.elapsed_time(e1) / hs
    composite = lv.nx * lv.ny + 2 * T.fine.nx * T.fine.ny
    # coupling kernels alone, on the final state
    lib, ctx, sp = L.load(), context(), C.c_void_p(st.cuda_stream)
    c, f = T.crse, T.fine
    reps = 20
    e0.record(st)
    for _ in range(reps):
        cp = c.patch(c.U_new)
        T._flux(cp, T.cfaces, 1, -1e-3 * c.dy, -1e-3 * c.dx, 0)
        for k in range(2):
            L.check(lib.hx_fill_fine_ghosts(ctx, C.byref(c.patch(c.U_old)), C.byref(cp), k / 2.0,
                                            2, box[0], box[1], C.byref(f.patch()), 0,
                                            f._ptr(f.U_new, 0), f._ptr(f.U_new, 1),
                                            f._ptr(f.U_new, 2), sp), "fill")
            fp = f.patch(f.U_new)
            T._flux(fp, T.ffaces, 2, 5e-4 * f.dy, 5e-4 * f.dx, 1)
        L.check(lib.hx_average_down_state(ctx, C.byref(f.patch()), 2, box[0], box[1],
                                          C.byref(cp), c.U_new[0].data_ptr(),
                                          c.U_new[1].data_ptr(), c.U_new[2].data_ptr(), sp),
                "average_down")
    e1.record(st)
    torch.cuda.synchronize()
    coupling_us = e0.elapsed_time(e1) / reps * 1e3
    set_arith("exact")
    out = {"workload": "two-level dam break: 4096^2 FP64 coarse, 1024^2-cell box at ratio 2 "
                       "(2048^2 fine), plain Roe, subcycled, CFL per level each step",
           "arith": arith, "steps": steps, "composite_cells_per_step": composite,
           "ms_per_step": ms, "wall_ms_per_step": wall / steps * 1e3,
           "composite_Mcups": composite / (ms / 1e3) / 1e6,
           "host_stepped_ms_per_step": host_ms,
           "coupling_kernels_us_per_step": coupling_us,
           "coupling_share": coupling_us / 1e3 / ms}
    del T, lv
    torch.cuda.empty_cache()
    return out


def time_amr_c3(steps, arith="fast", bed_amp=0.05):
    """Config C3 on an AMR hierarchy (workloads.circular_dam_break): 512^2
    base on [0,1]^2, depth 2.0 inside r = 0.2 around the centre and 0.5
    outside, bed bed_amp U[0,1) (SplitMix64; 0.05 is the config, whose noisy
    bed soon puts |grad h| above the threshold everywhere the water moves, so
    the fine levels come to cover the domain; 0 gives a flat bed on which
    they follow the front), Manning 0.03, hydrostatic Roe,
    transmissive sides, two refined levels at ratio 2, tags where
    |grad h| > 0.1, regrid every 4 steps of a level, blocking factor 16,
    subcycled, FP64.  Times `steps` coarse steps (after 4 warm-up steps) on
    the wall clock in calls of 2 coarse steps (level 1 regrids level 2 every
    2 coarse steps), each ending with a synchronisation, and counts composite
    cell updates (every level's cells times its substeps) on the grids at the
    start of each call."""
    import torch
    from hydroexa_amd import set_arith
    from hydroexa_amd import workloads as W
    from hydroexa_amd.amr import AmrHierarchy

    set_arith(arith)
    w = W.circular_dam_break(512, bed_amp=bed_amp)
    H = AmrHierarchy(w.solver, w.h, w.hu, w.hv, w.z, w.n, w.dx, w.dy, w.bc_lo, w.bc_hi,
                     max_level=2, regrid_int=4, blocking_factor=16, max_grid_size=128,
                     tag_thresh=[(0.1, 1e300, 1e300)])
    H.run_native(1e30, w.cfl, max_steps=4)
    torch.cuda.synchronize()
    cells = 0
    wall = 0.0
    done = 0
    boxes_seen = []
    while done < steps:
        k = min(2, steps - done)
        per = sum(lv.interior(lv.mask).sum().item() * (2 ** l if H.subcycle else 1)
                  for l, lv in enumerate(H.levels))
        boxes_seen.append([len(lv.boxes) for lv in H.levels])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        H.run_native(1e30, w.cfl, max_steps=k)
        torch.cuda.synchronize()
        wall += time.perf_counter() - t0
        cells += per * k
        done += k
    set_arith("exact")
    bed = f"bed {bed_amp:g} U[0,1)" if bed_amp else "flat bed"
    out = {"workload": "circular dam break (config C3): 512^2 base, max_level 2 at ratio 2, "
                       f"depth 2.0 / 0.5, {bed}, hydrostatic Roe + Manning 0.03, "
                       "transmissive, tags |grad h| > 0.1, regrid_int 4, subcycled, FP64",
           "arith": arith, "coarse_steps": steps, "wall_ms_per_coarse_step": wall / steps * 1e3,
           "composite_cell_updates_per_step": cells / steps,
           "composite_Mcups": cells / wall / 1e6,
           "boxes_per_level_first": boxes_seen[0] if boxes_seen else None,
           "boxes_per_level": [len(lv.boxes) for lv in H.levels],
           "cells_per_level": [int(lv.interior(lv.mask).sum().item()) for lv in H.levels],
           "time": H.time}
    del H
    torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.cpu_worker:
        _cpu_worker(int(args.cpu_worker[0]), float(args.cpu_worker[1]))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design "
                                   "note only (halo exchange + 8-byte dt min all-reduce)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    from hydroexa_amd import workloads as W

    if args.only == "ensemble":
        print(json.dumps({"ensemble": time_ensemble(args.ensemble_steps, args.arith)}), flush=True)
        return
    if args.only == "adjoint":
        print(json.dumps({"adjoint": time_adjoint(args.adjoint_steps, 20, args.arith)}), flush=True)
        return
    if args.only == "adjoint_ensemble":
        print(json.dumps({"adjoint_ensemble": time_adjoint_ensemble(args.adjoint_ensemble_steps,
                                                                    arith=args.arith)}), flush=True)
        return
    if args.only == "amr":
        print(json.dumps({"amr_c3": time_amr_c3(args.amr_steps, args.arith),
                          "amr_c3_flat_bed": time_amr_c3(args.amr_steps, args.arith, 0.0)}),
              flush=True)
        return
    if args.only == "two_level":
        print(json.dumps({"two_level": time_two_level(args.two_level_steps, args.arith)}),
              flush=True)
        return
    if args.solver == "Roe":
        w = W.dam_break(args.n, args.steps)
    else:
        w = W.lake_at_rest(args.n, args.steps)
    es = 8 if args.dtype == "f64" else 4
    ev_ms, wall, dth, fin = time_gpu(w, args.dtype, args.steps, args.warmup, args.rows,
                                     args.arith, keep=True, condition_ms=args.condition_ms)
    ms_per_step = ev_ms / args.steps
    cups = w.cells * args.steps / (ev_ms / 1e3)
    bytes_per_launch = 8 * es * w.cells          # read h,hu,hv,z,n; write h,hu,hv
    achieved = bytes_per_launch / (ms_per_step / 1e3) / 1e9
    # the other arithmetic build on the same workload, and the distance
    # between the two final states
    other = "exact" if args.arith == "fast" else "fast"
    ev2, _, _, fin2 = time_gpu(w, args.dtype, args.steps, args.warmup, args.rows, other, keep=True,
                               condition_ms=args.condition_ms)
    linf = [float((fin[k] - fin2[k]).abs().max() / fin2[k].abs().max().clamp_min(1e-300))
            for k in range(3)]
    kname = f"advance_{w.solver.lower()}_{args.dtype}_{args.arith}_{args.n}"
    traffic = traffic_from_profile(kname)
    out = {
        "metric": "Mcell-updates/s/GPU + %HBM roofline",
        "value": cups / 1e6,
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {"workload": w.name if args.dtype == "f64" else w.name.replace("fp64", "fp32"),
                   "cells": w.cells, "solver": w.solver, "cfl": w.cfl, "arith": args.arith,
                   "step": "fused advance with device CFL reduction", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kname,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "avg_launch_ms": ms_per_step},
        "arith_check": {"other": other, "other_ms_per_step": ev2 / args.steps,
                        "other_Mcups": w.cells * args.steps / (ev2 / 1e3) / 1e6,
                        "rel_linf_h_hu_hv": linf},
        "timing": {"warmup_steps": args.warmup, "timed_steps": args.steps,
                   "device_conditioning_ms": args.condition_ms,
                   "conditioning": "hx_stream_probe on its own buffers right before the warm-up "
                                   "steps, after the level's upload (the device's clock state "
                                   "after idle differs from its sustained one for ~30 steps)"},
        "wall_s": wall,
        "final_dt": float(dth[-1]) if len(dth) else None,
    }
    if args.dtype == "f64":
        ceil = stream_ceiling(w.cells)
        ceil["advance_over_ceiling"] = achieved / ceil["GBps"]
        out["roofline"]["stream_ceiling"] = ceil
    out["diagnostics"] = time_conservation(w, args.dtype)
    if args.extra:
        extra = {}
        bumpy = W.dam_break(4096, 1)
        rng = np.random.default_rng(5)
        bumpy.z = 0.05 * rng.uniform(size=bumpy.h.shape)
        bumpy.h = np.maximum(np.where(bumpy.h > 0.75, 1.0, 0.5) - bumpy.z, 0.0)
        bumpy.n = np.full_like(bumpy.h, 0.03)
        cases = (("dam_break_4096_bumpy_rough_f64_fast", bumpy, "f64", "fast", 0),
                 ("dam_break_4096_bumpy_rough_f64_exact", bumpy, "f64", "exact", 0),
                 ("dam_break_4096_f32_fast", W.dam_break(4096, 1), "f32", "fast", 0),
                 ("lake_at_rest_1024_f64_fast", W.lake_at_rest(1024, 1), "f64", "fast", 0),
                 ("lake_at_rest_1024_f64_exact", W.lake_at_rest(1024, 1), "f64", "exact", 0),
                 ("lake_at_rest_2048_f64_fast", W.lake_at_rest(2048, 1), "f64", "fast", 0),
                 ("lake_at_rest_1024_f32_fast", W.lake_at_rest(1024, 1), "f32", "fast", 0))
        for name, wk, dt_, ar, rows in cases:
            ms, _, _, _ = time_gpu(wk, dt_, 200, 10, rows, ar, condition_ms=args.condition_ms)
            e = 8 if dt_ == "f64" else 4
            extra[name] = {"Mcups": wk.cells * 200 / (ms / 1e3) / 1e6, "ms_per_step": ms / 200,
                           "hbm_frac": 8 * e * wk.cells / (ms / 200 / 1e3) / 1e9 / HBM_PEAK_GBS}
        out["extra"] = extra
    if args.ensemble_steps > 0:
        out["ensemble"] = time_ensemble(args.ensemble_steps, args.arith)
    if args.adjoint_steps > 0:
        out["adjoint"] = time_adjoint(args.adjoint_steps, 20, args.arith)
    if args.adjoint_ensemble_steps > 0:
        out["adjoint_ensemble"] = time_adjoint_ensemble(args.adjoint_ensemble_steps, arith=args.arith)
    if args.two_level_steps > 0:
        out["two_level"] = time_two_level(args.two_level_steps, args.arith)
    if args.amr_steps > 0:
        out["amr_c3"] = time_amr_c3(args.amr_steps, args.arith)
        out["amr_c3_flat_bed"] = time_amr_c3(args.amr_steps, args.arith, 0.0)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.n, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
