#!/usr/bin/env python3
"""Throughput of the batched steady-state J-V solve on one MI355X.

A step is one `solve_batch` of the whole synthetic workload: 16384 devices (NP = 400) with
mobile anions and cations in the perovskite, free at every voltage, 61 voltages from -0.2 to
1.3 V, in mixed precision (FP32 block factorisation of each voltage's first Newton step on
row-equilibrated blocks, every step's factors stored in FP32; FP64 residual, exponentials,
back-substitution, current and stop rules). The all-FP64
form is timed beside it, and the mixed curves are checked against it (1e-6 of each curve's
scale). One JSON line on stdout:

    python bench.py [--steps K] [--warmup W] [--n 16384] [--precision mixed|fp64]

``roofline``: the solve kernel's algorithmic HBM bytes (per needed Newton step per device
B_step = NP*(16b + 8s + 2w(b^2 + bk)), b = 3, s = 22, k = 3 with the ion border, w the
factor's bytes) over its duration from HIP events on its launch stream, against 8 TB/s.
``cpu_baseline``: the FP64 CPU oracle (a C restatement of the reference's algorithm) timed
on this host, one pinned single-threaded process per core of the GPU's CPU share, each on its
own devices of the same workload, as the reference times its CPU path; ``vs_oracle`` holds
the GPU curves of those devices against the oracle's (mixed, and FP64 under ``fp64``).
``lockstep``: Newton steps needed against steps paid (each wavefront runs until its slowest
device is done). ``frozen``: the same workload with the ions frozen after a 1.2 V pre-bias,
timed and checked against the oracle. ``gradient`` (dJ/dtheta of 4096 devices without ions)
and ``impedance`` (2048 devices, NP = 300, 50 frequencies) are legs of their own.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "J-V curves/s, 16384 devices NP=400, mixed vs FP64 at <=1e-6 J-scale error"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
SEGMENT_COEFFS = 22  # s: per-point coefficients of the reference's stencils
CHAOTIC_STEPS = 20


def b_step(NP: int, k: int, w: int, b: int = 3, s: int = SEGMENT_COEFFS) -> int:
    """Algorithmic bytes per Newton step per device: state read+write, coefficients once,
    factors written by the elimination and read by the back-substitution."""
    return NP * (16 * b + 8 * s + 2 * w * (b * b + b * k))


def parse():
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--n", type=int, default=16384)
    p.add_argument("--config", default="c3", choices=["c2", "c3"])
    p.add_argument("--variant", default="free", choices=["free", "frozen"])
    p.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    p.add_argument("--lanes", type=int, default=2, choices=[1, 2, 4],
                   help="GPU lanes per device (2: each Newton solve split at the grid's middle;"
                        " 4, in quarters: c2 only, the devices without mobile ions)")
    p.add_argument("--grad-lanes", type=int, default=4, choices=[1, 2, 4],
                   help="GPU lanes per device of the gradient workload's solve")
    p.add_argument("--grad-precision", default="fp64", choices=["mixed", "fp64"],
                   help="precision of the gradient workload's solve (the adjoint is FP64)")
    p.add_argument("--no-fp64", action="store_true", help="skip the all-FP64 comparison")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--cpu-curves", type=int, default=24, help="curves per CPU process")
    p.add_argument("--workers", type=int, default=0, help="host processes building devices")
    p.add_argument("--grad-n", type=int, default=4096,
                   help="devices of the gradient workload (dJ/dtheta, 16 parameters); 0: skip")
    p.add_argument("--frozen-n", type=int, default=16384,
                   help="devices of the c3 frozen-after-pre-bias leg (timed beside the free "
                        "one); 0: skip")
    p.add_argument("--imp-n", type=int, default=2048,
                   help="devices of the impedance workload (NP=300, mobile ions); 0: skip")
    p.add_argument("--imp-freqs", type=int, default=50,
                   help="frequencies of the impedance workload (log-spaced, 1e-2..1e7 Hz)")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="only copy MIB MiB device to device (HBM counter calibration)")
    return p.parse_args()


#: The host cores one GPU's share of the box gives a job (the CPU leg's processes and the
#: host builders' pools stay within it whatever the affinity mask shows).
CPU_SHARE = 16


def affinity_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def host_cores() -> int:
    return max(1, min(CPU_SHARE, affinity_cores()))


# ---- CPU baseline: the oracle, one pinned single-threaded process per core -----------------
def _cpu_worker(core, rows, config, variant, start_evt, ready, out):
    os.environ["OMP_NUM_THREADS"] = "1"
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    try:
        os.sched_setaffinity(0, {core})
    except (AttributeError, OSError):
        pass
    from jaxdd_amd import synthetic as S
    from oracle import ddoracle as O

    O.lib()
    t = S.draw(rows[1], S.SEED)
    ions = config == "c3"
    devs = [S.device_from_row(t, k, 400, ions, ions and variant == "frozen")
            for k in range(*rows)]
    O.solve_batch(devs[:1], S.VOLTAGES)  # warm-up
    ready.put(core)
    start_evt.wait()
    t0 = time.perf_counter()
    J, ok, its = O.solve_batch(devs, S.VOLTAGES)
    t1 = time.perf_counter()
    out.put((t0, t1, len(devs), float(ok.mean()), rows, J, ok, its))


def cpu_baseline(config, variant, per_proc):
    import multiprocessing as mp

    ctx = mp.get_context("fork")
    cores = host_cores()
    try:
        core_ids = sorted(os.sched_getaffinity(0))[:cores]
    except AttributeError:
        core_ids = list(range(cores))
    start, ready, out = ctx.Event(), ctx.Queue(), ctx.Queue()
    procs = [ctx.Process(target=_cpu_worker,
                         args=(c, (i * per_proc, (i + 1) * per_proc), config, variant, start,
                               ready, out))
             for i, c in enumerate(core_ids)]
    for p in procs:
        p.start()
    for _ in procs:  # every worker has built its devices and warmed up
        ready.get(timeout=600)
    start.set()
    res = [out.get(timeout=600) for _ in procs]
    for p in procs:
        p.join()
    t0 = min(r[0] for r in res)
    t1 = max(r[1] for r in res)
    curves = sum(r[2] for r in res)
    res.sort(key=lambda r: r[4][0])
    oracle_curves = tuple(np.concatenate([r[k] for r in res]) for k in (5, 6, 7))
    return {"value": curves / (t1 - t0), "unit": "J-V curves/s", "cores": len(procs),
            "affinity_cores": affinity_cores(), "kind": "port",
            "sample": f"{curves} curves ({per_proc} devices per process, devices 0-{curves - 1}) "
                      f"of the same workload, FP64 C oracle with the reference's two passes, "
                      f"{len(procs)} pinned single-threaded processes (the box's CPU share for "
                      "one GPU)",
            "per_core": curves / (t1 - t0) / len(procs),
            "converged": float(np.mean([r[3] for r in res]))}, oracle_curves


def versus_oracle(J, ok, its, oracle_curves, J_scale, wild=None):
    """The GPU's curves of the devices the CPU leg solved against the oracle's: the largest
    |J - J_oracle| over each curve's scale (J_scale + max |J_oracle|) where both converged,
    converged flags equal where neither side took CHAOTIC_STEPS steps or more, and the points
    whose Newton steps differ."""
    Jo, oko, itso = oracle_curves
    m = len(Jo)
    J, ok, its = J[:m], ok[:m], its[:m]
    keep = np.ones(m, bool) if wild is None else ~wild
    excluded = np.flatnonzero(~keep).tolist()
    J, ok, its, Jo, oko, itso = (a[keep] for a in (J, ok, its, Jo, oko, itso))
    scale = J_scale[:m][keep, None] + np.max(np.abs(Jo), axis=1, keepdims=True)
    both = ok & oko
    calm = (its < CHAOTIC_STEPS) & (itso < CHAOTIC_STEPS)
    return {"devices": m, "excluded_chaotic_pre_bias": excluded,
            "max_err": float(np.max((np.abs(J - Jo) / scale)[both], initial=0.0)),
            "flags_equal_calm": bool(np.array_equal(ok[calm], oko[calm])),
            "chaotic_points": int((~calm).sum()),
            "steps_differ": int(np.sum(its[calm] != itso[calm]))}


def lockstep(its, lanes: int):
    """Newton steps needed (the sum of every device's iterations) against steps paid: a
    wavefront (64 / lanes devices) runs until its slowest device has finished its scan, so
    it pays its devices' count times their largest total (the reference's lockstep,
    benchmarks/run.py:108-113, whose vmap pays the batch's slowest device at every voltage)."""
    per = its.sum(axis=1)
    w = 64 // lanes
    paid = sum(len(per[i:i + w]) * int(per[i:i + w].max()) for i in range(0, len(per), w))
    needed = int(per.sum())
    nV = its.shape[1]
    return {"steps_needed_per_voltage": needed / its.size,
            "steps_paid_per_voltage": paid / its.size, "lockstep": paid / needed,
            "voltages": nV}


def copy_calibration(mib: int) -> None:
    """A plain device-to-device copy of known size, for scaling rocprofv3's FETCH_SIZE and
    WRITE_SIZE counters to bytes."""
    import torch

    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "dispatches": 3}))


def adjoint_b_unit(NP: int = 400, b: int = 3, s: int = SEGMENT_COEFFS) -> int:
    """Algorithmic bytes of the adjoint per converged (device, voltage): the assembly at the
    converged state, the transposed sweep's factors written and read, the reverse pass's reads
    of lambda (its contraction with the 16 directions' tangents reads each device's tangents
    once for all its voltages)."""
    return NP * (8 * (b + s) + 2 * 8 * (b * b + b) + 8 * b)


def gradient_bench(theta, V, warm: bool = True, lanes: int = 4, precision: str = "fp64"):
    """dJ/dtheta of the gradient workload's curves from the parameter vectors: each device's
    graded layout, its packed arrays and their 16 parameter tangents built on the GPU
    (ddh_graded_layout, ddh_build), one solve keeping the converged states and one adjoint
    launch, timed with HIP events on the launch stream and by the host clock around the whole
    (a warm-up round has allocated the workspaces)."""
    import torch

    from jaxdd_amd.adjoint import gpu_tangents, jv_gradients

    n = len(theta)
    if warm:
        batch, tan = gpu_tangents(theta)
        jv_gradients(batch, tan, V, lanes=lanes, precision=precision)
        del batch, tan
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch, tan = gpu_tangents(theta)
    tb = time.perf_counter()
    J, ok, its, grad = jv_gradients(batch, tan, V, events=ev, lanes=lanes, precision=precision)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    solve_ms, adj_ms = ev[0].elapsed_time(ev[1]), ev[2].elapsed_time(ev[3])
    ok = ok.cpu().numpy()
    unit = adjoint_b_unit()
    achieved = float(ok.sum()) * unit / (adj_ms * 1e-3) / 1e9
    return {
        "value": n / (t1 - t0),
        "unit": "gradient curves/s (dJ(V)/dtheta, 16 parameters, from the parameter vectors: "
                "the layouts, the batch and its tangents built on the GPU, the solve and the "
                "adjoint)",
        "devices": n, "voltages": len(V), "ms": (t1 - t0) * 1e3,
        "build_ms": (tb - t0) * 1e3, "tangent_bytes": sum(v.numel() * 8 for v in tan.values()),
        "solve_kernel_ms": solve_ms, "adjoint_kernel_ms": adj_ms, "solve_lanes_per_device": lanes,
        "solve_precision": precision,
        "converged_frac": float(ok.mean()),
        "finite": bool(np.isfinite(grad.cpu().numpy()).all()),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "kernel": "ddh::adjoint_kernel",
                     "bytes_per_converged_voltage": unit},
    }


#: FP64 operations per (device, frequency, grid point) of ddh::impedance_kernel<false> (3 x 3,
#: no mobile ions), counted from its code (complex multiply 6, complex add 2, complex divide 11,
#: real x complex 2): the occupation's lag 12, the diagonal's time and trap terms 28, the
#: elimination's A C' and A d' with A's five entries 74, the block inverse by its adjugate 213,
#: d' 66, C' from C's five entries 42, the current's coefficients 19, the contraction's g and
#: its sum 96.
IMP_FLOPS_PER_POINT = 550
#: ... of ddh::impedance_kernel<true> (5 x 5 with the ions as unknowns, three right-hand sides,
#: three functionals), counted the same way: the occupation's lag 13, the diagonal's time and
#: trap terms 28, the current's coefficients 25, the elimination's A C' (A's nine entries) 190
#: and A d' 114, the functionals' g (3 x 5 x 5) 600, the partitioned inverse 934 (the 3 x 3
#: adjugate 213, P^-1 Q and R P^-1 264, the 2 x 2 Schur complement and its inverse 145, the
#: blocks of the inverse 312), d' = inv rhs 600, C' (column 0 dense, four diagonal entries)
#: 140, the functionals' sums (3 x 3 x 5) 360.
IMP_ION_FLOPS_PER_POINT = 3004


def impedance_bench(packed, freqs, voltage=0.0):
    """The impedance workload: the steady solve at ``voltage`` keeping its states, then
    ddh_impedance over every (device, frequency), the batch resident in HBM before the timed
    call (a warm-up call has run), timed with HIP events on the launch stream and by the host
    clock around the call."""
    import torch

    from jaxdd_amd.batch import SC_FIELDS, from_packed
    from jaxdd_amd.impedance import solve_impedance

    batch = from_packed(packed, "cuda")
    n, NP = batch["n"], batch["NP"]
    sc = packed["sc"]
    ions = bool(np.any(sc[SC_FIELDS.index("a_total")] > 0) or np.any(sc[SC_FIELDS.index("c_total")] > 0))
    flops = IMP_ION_FLOPS_PER_POINT if ions else IMP_FLOPS_PER_POINT
    solve_impedance(batch, freqs, voltage=voltage)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = solve_impedance(batch, freqs, voltage=voltage, events=ev)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    steady_ms, imp_ms = ev[0].elapsed_time(ev[1]), ev[2].elapsed_time(ev[3])
    systems = n * len(freqs)
    achieved = systems * NP * flops / (imp_ms * 1e-3) / 1e12
    return {
        "value": n / (t1 - t0),
        "unit": f"impedance spectra/s (steady state + {len(freqs)} frequencies)",
        "devices": n, "frequencies": len(freqs), "ms": (t1 - t0) * 1e3,
        "steady_kernel_ms": steady_ms, "impedance_kernels_ms": imp_ms,
        "systems_per_s": systems / (imp_ms * 1e-3),
        "converged_frac": float(r.converged.mean()), "finite": bool(np.isfinite(r.Y).all()),
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "kernel": f"ddh::impedance_prep_kernel<{str(ions).lower()}> + "
                               f"ddh::impedance_kernel<{str(ions).lower()}>",
                     "flops_per_point": flops},
    }


def traffic_from_profiles(workload: str):
    """HBM bytes per launch of the solve kernel from the committed rocprofv3 counter summary
    (profiles/hbm_traffic.json), when it was measured on this workload."""
    f = ROOT / "profiles" / "hbm_traffic.json"
    if not f.exists():
        return None
    d = json.loads(f.read_text())
    return d.get("bytes_per_launch") if d.get("workload") == workload else None


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the devices of a batch are independent, so a "
              "batch splits across GPUs by devices with no communication, and the throughput "
              f"measured is one GPU's share; --gpus {args.gpus} is not supported.",
              file=sys.stderr)
        return 2
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return 0
    from jaxdd_amd import synthetic as S

    workers = args.workers or host_cores()
    t = time.perf_counter()
    packed = S.packed(args.config, args.n, variant=args.variant, workers=workers)
    build_s = time.perf_counter() - t
    grad_inputs, grad_host_s = None, 0.0
    if args.grad_n:
        from jaxdd_amd import adjoint as A

        t = time.perf_counter()
        theta = A.theta_from_table(S.draw(args.grad_n), range(args.grad_n))
        grad_inputs = (theta,)
        grad_host_s = time.perf_counter() - t
    imp_packed = None
    if args.imp_n:
        # the c5 devices: the c3 stack with its mobile ions free, on 300 points
        imp_packed = S.packed("c3", args.imp_n, workers=workers, NP=300)
    frozen_packed = None
    if args.frozen_n and args.config == "c3" and args.variant == "free":
        frozen_packed = S.packed("c3", args.frozen_n, variant="frozen", workers=workers)
    cpu, oracle_curves = None, None
    if not args.no_cpu:
        cpu, oracle_curves = cpu_baseline(args.config, args.variant, args.cpu_curves)
    frozen_oracle = frozen_wild = None
    if frozen_packed is not None:
        from oracle import ddoracle as O

        import dataclasses

        fdevs = S.devices("c3", 32, variant="frozen")
        frozen_oracle = O.solve_batch(fdevs, S.VOLTAGES)
        # devices whose pre-bias lands on one of several nearby profiles by the last bits of
        # the arithmetic: the oracle's own curve moves by more than 1e-7 of its scale when the
        # generation rate changes by 1e-13
        moved = O.solve_batch([dataclasses.replace(d, G=d.G * (1 + 1e-13)) for d in fdevs],
                              S.VOLTAGES)[0]
        fscale = (np.array([O.j_scale(d) for d in fdevs])[:, None]
                  + np.max(np.abs(frozen_oracle[0]), axis=1, keepdims=True))
        frozen_wild = np.max(np.abs(moved - frozen_oracle[0]) / fscale, axis=1) > 1e-7

    import torch

    from jaxdd_amd.batch import from_packed, solve_batch

    batch = from_packed(packed, "cuda")
    V = S.VOLTAGES
    for _ in range(args.warmup):
        solve_batch(batch, V, precision=args.precision, lanes=args.lanes)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        J, ok, its = solve_batch(batch, V, precision=args.precision, events=evs[k],
                                 lanes=args.lanes)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    kernel_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    J, ok, its = J.cpu().numpy(), ok.cpu().numpy(), its.cpu().numpy()

    ions = args.config == "c3"
    w = 4 if args.precision == "mixed" else 8
    step_bytes = b_step(400, 3 if ions else 1, w)
    launch_bytes = float(its.sum()) * step_bytes
    achieved = launch_bytes / (kernel_ms * 1e-3) / 1e9
    workload = f"{args.config}-{args.variant}-n{args.n}-{args.precision}-lanes{args.lanes}"
    kernel = {1: "ddh::solve_kernel", 2: "ddh::solve_pair_kernel", 4: "ddh::solve_quad_kernel"}[args.lanes]

    out = {
        "metric": METRIC,
        "value": args.n * args.steps / elapsed,
        "unit": "J-V curves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.config} {'free' if args.variant == 'free' else 'frozen-after-1.2V'}"
                         f"-ion J-V: {args.n} devices, NP=400, 61 voltages -0.2..1.3 V, "
                         f"3x3 blocks{' + 2-slot ion border' if ions else ''}, "
                         f"{args.precision} precision, seed {S.SEED}"),
            "devices": args.n, "NP": 400, "voltages": len(V), "precision": args.precision,
            "lanes_per_device": args.lanes,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profiles(workload),
            "kernel": kernel, "kernel_ms": kernel_ms,
            "algorithmic_bytes_per_launch": launch_bytes, "B_step": step_bytes,
        },
        "cpu_baseline": cpu,
        "converged_frac": float(ok.mean()),
        "newton_steps_per_voltage": float(its.mean()),
        "lockstep": lockstep(its, args.lanes),
        "host_build_s": build_s,
    }
    J_scale = batch["J_scale"].cpu().numpy()
    if oracle_curves is not None:
        out["vs_oracle"] = versus_oracle(J, ok, its, oracle_curves, J_scale)
    if not args.no_fp64 and args.precision == "mixed":
        e64 = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        torch.cuda.synchronize()
        a = time.perf_counter()
        J64, ok64, its64 = solve_batch(batch, V, precision="fp64", events=e64, lanes=args.lanes)
        torch.cuda.synchronize()
        b = time.perf_counter()
        J64, ok64, its64 = J64.cpu().numpy(), ok64.cpu().numpy(), its64.cpu().numpy()
        scale = batch["J_scale"].cpu().numpy()[:, None] + np.max(np.abs(J64), axis=1, keepdims=True)
        both = ok & ok64
        calm = (its < CHAOTIC_STEPS) & (its64 < CHAOTIC_STEPS)
        out["fp64"] = {
            "value": args.n / (b - a), "ms_per_step": (b - a) * 1e3,
            "kernel_ms": e64[0].elapsed_time(e64[1]),
            "newton_steps_per_voltage": float(its64.mean()),
            "speedup_mixed_over_fp64": (b - a) / (elapsed / args.steps),
        }
        out["max_err_vs_fp64"] = float(np.max((np.abs(J - J64) / scale)[both]))
        out["flags_equal_calm"] = bool(np.array_equal(ok[calm], ok64[calm]))
        out["chaotic_points"] = int((~calm).sum())
        if oracle_curves is not None:
            out["fp64"]["vs_oracle"] = versus_oracle(J64, ok64, its64, oracle_curves, J_scale)
    if frozen_packed is not None:
        fb = from_packed(frozen_packed, "cuda")
        solve_batch(fb, V, precision=args.precision, lanes=args.lanes)  # warm-up
        ef = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        torch.cuda.synchronize()
        a = time.perf_counter()
        Jf, okf, itsf = solve_batch(fb, V, precision=args.precision, events=ef, lanes=args.lanes)
        torch.cuda.synchronize()
        b = time.perf_counter()
        Jf, okf, itsf = Jf.cpu().numpy(), okf.cpu().numpy(), itsf.cpu().numpy()
        fkernel = ef[0].elapsed_time(ef[1])
        fbytes = float(itsf.sum()) * step_bytes
        out["frozen"] = {
            "value": args.frozen_n / (b - a), "unit": "J-V curves/s", "ms_per_step": (b - a) * 1e3,
            "kernel_ms": fkernel, "workload": (f"c3 frozen after a 1.2 V pre-bias: "
                                               f"{args.frozen_n} devices, NP=400, 61 voltages, "
                                               f"{args.precision} precision"),
            "converged_frac": float(okf.mean()), "newton_steps_per_voltage": float(itsf.mean()),
            "lockstep": lockstep(itsf, args.lanes),
            "roofline": {"bound": "hbm", "achieved": fbytes / (fkernel * 1e-3) / 1e9,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": fbytes / (fkernel * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "kernel": kernel, "B_step": step_bytes},
            "vs_oracle": versus_oracle(Jf, okf, itsf, frozen_oracle,
                                       fb["J_scale"].cpu().numpy(), frozen_wild),
        }
        del fb
    if grad_inputs is not None:
        out["gradient"] = gradient_bench(*grad_inputs, V, lanes=args.grad_lanes,
                                         precision=args.grad_precision)
        g = out["gradient"]
        g["host_draw_s"] = grad_host_s
        # the whole gradient, the parameter vectors' draw on the host included
        g["end_to_end_value"] = args.grad_n / (grad_host_s + g["ms"] * 1e-3)
        out["gradient"]["workload"] = (f"c4: {args.grad_n} devices without ions, NP=400, 61 "
                                       "voltages, FP64, 16 parameters")
    if imp_packed is not None:
        freqs = np.logspace(-2, 7, args.imp_freqs)
        out["impedance"] = impedance_bench(imp_packed, freqs)
        out["impedance"]["workload"] = (f"c5: {args.imp_n} devices with mobile ions, NP=300, "
                                        f"{args.imp_freqs} frequencies 1e-2..1e7 Hz at 0 V, "
                                        "complex128, 5x5 blocks with the slots' border")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
