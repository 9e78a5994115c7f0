"""Benchmark of the explicit finite-volume step on one MI355X.

Headline workload: the 2-D Kapila five-equation two-phase model, MUSCL with
the van Leer limiter, FP64, 4096 x 4096 uniform Cartesian cells: a planar
Mach 2.4 air shock hitting 32 water cylinders (radii uniform in [20, 80]
cells, seed-placed without overlap), alpha floor 1e-6.  Synthetic initial
state, built on the host and uploaded before timing; the timed region runs
--steps full time steps (predictor + corrector kernels each) with the state
resident in HBM.

Prints one JSON line: cell-updates/s, the HBM roofline of the dominant kernel
measured with HIP events over the timed region, a CPU baseline (the
reference executable built from source, MPI ranks on the host cores, or the
FP64 oracle restatement) on a bounded sample of the same workload,
and the secondary measurements: FP32 storage on the same workload, the
batched pool (2048 water/air Riemann problems, FP32 storage) against serial
plain FP64 solves, the served request stream at three arrival rates, the
1-D first-order shock-tube latency and the 3-D pressure-disequilibrium
bubble workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "cell-updates/s & % HBM roofline; batched solves/s + p99 latency vs plain"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=4096, help="cells per side")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-kind", choices=["auto", "reference", "port"], default="auto",
                    help="CPU baseline: the reference executable (auto: when built) or the oracle restatement")
    ap.add_argument("--cpu-n", type=int, default=2048, help="cells per side of the oracle (port) sample")
    ap.add_argument("--cpu-steps", type=int, default=20, help="steps of the oracle (port) sample")
    ap.add_argument("--ref-n", type=int, default=1024, help="cells per side of the reference sample")
    ap.add_argument("--ref-steps", type=int, default=10, help="measured steps of the reference sample")
    ap.add_argument("--no-fp32s", action="store_true", help="skip the FP32-storage measurement")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain copy kernel of known size (for HBM counter calibration)")
    ap.add_argument("--batch-cases", type=int, default=2048, help="cases of the batched leg (0: skip)")
    ap.add_argument("--serial-stride", type=int, default=32,
                    help="serial plain FP64 comparator: every k-th case of the batch")
    ap.add_argument("--serve-requests", type=int, default=120, help="requests per arrival rate (0: skip)")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[50.0, 200.0, 800.0])
    ap.add_argument("--c3-n", type=int, default=384)
    ap.add_argument("--c3-steps", type=int, default=20, help="steps of the 3-D leg (0: skip)")
    ap.add_argument("--only-main", action="store_true", help="headline workload only")
    ap.add_argument("--legs", nargs="+", choices=["latency", "batched", "served", "c3"],
                    default=["latency", "batched", "served", "c3"], help="secondary measurements to run")
    ap.add_argument("--batch-direct", action="store_true",
                    help="batched leg with direct launches instead of hipGraph replay (for kernel tracing)")
    ap.add_argument("--no-serial", action="store_true", help="batched leg without the serial FP64 comparator")
    return ap.parse_args()


def log(msg):
    """Progress on stderr (stdout carries only the JSON line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def state_words(case):
    nph = len(case["eos"])
    dim = sum(1 for c in case["cells"] if c > 1)
    return (3 * nph + dim + 1) if case["model"] != "euler" else dim + 2


def copy_calibration(nbytes=1 << 30):
    """Plain copy of nbytes through the library's streaming copy kernel."""
    from ecogen_amd import lib
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    hip.hipMemset(a, 0, C.c_size_t(nbytes))
    for _ in range(3):
        assert lib().eco_copy_calibration(b, a, nbytes, None) == 0
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    return nbytes


REF_BIN = ROOT / "oracle" / "_ref" / "ECOGEN"   # the reference, built from its sources by oracle/ref.mk
MPIEXEC = Path("/opt/conda/bin/mpiexec")         # MPICH of the image, which the reference links


def _ref_deck_case(n, steps):
    """Deck of the headline workload at n x n cells (cylinder radii scaled)."""
    from ecogen_amd import workloads as W
    sys.path.insert(0, str(ROOT / "oracle"))
    from deck import OUTPUT_BINARY
    scale = n / 4096.0
    air, shocked, water = W.shock_cylinders_2d_states()
    doms = [{"type": "entireDomain", "state": "air"},
            {"type": "halfSpace", "state": "shocked", "axe": "x", "origin": 0.05, "direction": "negative"}]
    doms += [{"type": "disc", "state": "water", "radius": r, "center": c}
             for c, r in W.cylinder_bodies(n, 32, 0xEC0C0002, 20.0 * scale, 80.0 * scale)]
    return {"name": "bench", "model": "kapila", "alpha_null": False, "eos": [W.AIR, W.WATER],
            "cells": [n, n, 1], "lengths": [1.0, 1.0, 1.0], "order": 2, "limiters": ["vanleer"] * 4,
            "cfl": 0.8, "time_mode": "iterations", "steps": steps, "output": OUTPUT_BINARY,
            "domains": doms, "states": [dict(air, u=air["u"][:2]), dict(shocked, u=shocked["u"][:2]),
                                        dict(water, u=water["u"][:2])]}


def _run_reference(n, steps, ranks):
    import shutil
    import subprocess
    import tempfile
    sys.path.insert(0, str(ROOT / "oracle"))
    from deck import write_deck
    tmp = Path(tempfile.mkdtemp(prefix="eco_ref_"))
    try:
        write_deck(_ref_deck_case(n, steps), tmp)
        t0 = time.perf_counter()
        subprocess.run([str(MPIEXEC), "-n", str(ranks), str(REF_BIN)], cwd=tmp, capture_output=True, text=True,
                       timeout=900, check=True)
        return time.perf_counter() - t0
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def cpu_baseline(args):
    """The reference's own CPU path (its executable built from source, MPI ranks
    on the host cores) when it is present, else the FP64 oracle restatement
    (OpenMP).  The reference is timed at two step counts so that start-up,
    mesh construction and output cancel: rate = cells x (s2 - s1) / (T2 - T1)."""
    n = args.cpu_n
    # one GPU's share of the host is 16 cores (the GPU host gives each GPU's
    # job 16); the host's total count is reported beside it
    cores = min(16, os.cpu_count() or 1)
    have_ref = REF_BIN.exists() and MPIEXEC.exists()
    if args.cpu_kind == "reference" or (args.cpu_kind == "auto" and have_ref):
        n_ref = args.ref_n
        s1, s2 = 2, 2 + args.ref_steps
        t1 = _run_reference(n_ref, s1, cores)
        t2 = _run_reference(n_ref, s2, cores)
        return {"value": n_ref * n_ref * (s2 - s1) / (t2 - t1), "unit": "cell-updates/s", "cores": cores,
                "host_cores": os.cpu_count(), "kind": "reference",
                "sample": f"headline workload at {n_ref}x{n_ref} cells (cylinder radii scaled), reference ECOGEN "
                          f"built from source, mpiexec -n {cores} (MPICH 3.3.2); {s2} minus {s1} steps, wall "
                          f"{t2:.1f} s - {t1:.1f} s"}
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    from ecogen_amd import workloads as W
    scale = n / 4096.0
    case, st = W.shock_cylinders_2d(n=n, rmin=20.0 * scale, rmax=80.0 * scale, steps=args.cpu_steps)
    t0 = time.perf_counter()
    _, it, _, _, _ = O.run(case, st, nsteps=args.cpu_steps, nthreads=cores)
    dt = time.perf_counter() - t0
    return {"value": n * n * it / dt, "unit": "cell-updates/s", "cores": cores, "host_cores": os.cpu_count(),
            "kind": "port",
            "sample": f"same workload scaled to {n}x{n} cells (cylinder radii scaled), {it} steps, FP64 oracle "
                      f"restatement of the reference step, OpenMP {cores} threads, {dt:.1f} s"}


def traffic_from_profiles(n, precision, kernel):
    """HBM bytes per launch of a stage kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (scaled by the copy-kernel calibration), or
    None when no profile of this workload size is committed."""
    p = ROOT / "profiles" / f"traffic_kapila2d_o2_vanleer_{n}.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text())["kernels"][precision][kernel]["bytes_per_launch"]
    except (KeyError, ValueError):
        return None


def valu_from_profiles(n, precision, kernel, kernel_ms):
    """VALU instructions per launch of a stage kernel from the committed
    rocprofv3 SQ_INSTS_VALU pass, and the fraction of the VALU issue capacity
    they fill in the live kernel time (an instruction of a 64-wide wave takes
    4 cycles of a 16-lane SIMD; 1024 SIMDs at 2.4 GHz), or None."""
    p = ROOT / "profiles" / f"valu_kapila2d_{n}.json"
    if not p.exists():
        return None
    try:
        insts = json.loads(p.read_text())["kernels"][f"{precision} {kernel}"]["valu_per_launch"]
    except (KeyError, ValueError):
        return None
    return {"insts_per_launch": int(insts), "issue_frac": round(insts * 4 / (1024 * 2.4e9 * kernel_ms * 1e-3), 4),
            "source": p.name}


def run_precision(case, st, args, precision):
    """Upload, warm up and time ``args.steps`` steps; returns the measurements."""
    from ecogen_amd import Context
    c = dict(case, precision=precision)
    ncell = int(np.prod(c["cells"]))
    ctx = Context(c)
    ctx.set_state(st)
    if args.warmup > 0:
        ctx.run(nsteps=args.warmup)
    ctx.sync()
    t0 = time.perf_counter()
    status, stage_ms, launches = ctx.time_steps(args.steps)
    t1 = time.perf_counter()
    if status.flags & 2:
        raise SystemExit(f"non-finite state produced ({precision})")
    out = ctx.get_state()
    ctx.close()
    wall = t1 - t0
    return {"value": ncell * args.steps / wall, "wall": wall, "stage_ms": stage_ms, "launches": launches,
            "status": status, "state": out}


def roofline(case, r, word_bytes):
    """Dominant stage kernel: algorithmic bytes (predictor reads U^n, writes U*:
    2 S words per cell; corrector reads U^n and U*, writes U^n+1: 3 S words)
    over its mean HIP-event duration in the timed region."""
    ncell = int(np.prod(case["cells"]))
    sw = state_words(case) * word_bytes
    per_cell = {0: 2 * sw, 1: 2 * sw, 2: 3 * sw}
    names = {0: "first-order step", 1: "MUSCL predictor", 2: "MUSCL corrector"}
    ms, cnt = r["stage_ms"], r["launches"]
    dom = max(range(3), key=lambda s: ms[s])
    avg_s = ms[dom] / max(cnt[dom], 1) / 1e3
    alg = per_cell[dom] * ncell
    ach = alg / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "kernel": names[dom], "kernel_ms": round(avg_s * 1e3, 4),
            "alg_bytes_per_launch": alg,
            "step_frac": round(5 * sw * r["value"] / (HBM_PEAK_GBS * 1e9), 4)}


def l1_errors(case, a, b):
    nph = len(case["eos"])
    dim = sum(1 for c in case["cells"] if c > 1)
    q = {}
    for name, fa, fb in (("rho", sum(a[3 * k] * a[3 * k + 1] for k in range(nph)),
                                 sum(b[3 * k] * b[3 * k + 1] for k in range(nph))),
                         ("p", sum(a[3 * k] * a[3 * k + 2] for k in range(nph)),
                               sum(b[3 * k] * b[3 * k + 2] for k in range(nph))),
                         ("velocity", np.concatenate([a[3 * nph + d] for d in range(dim)]),
                                      np.concatenate([b[3 * nph + d] for d in range(dim)]))):
        q[name] = float(np.abs(fa - fb).sum() / np.abs(fb).sum())
    for k in range(nph):
        q[f"alpha{k}"] = float(np.abs(a[3 * k] - b[3 * k]).sum() / np.abs(b[3 * k]).sum())
    return q


def _pct(x, q):
    return float(np.percentile(np.asarray(x), q)) if len(x) else None


def _fp32s_vs_fp64(case, a32, a64):
    """rel L1 of FP32 storage against FP64 in mixture density, mixture pressure,
    velocity and volume fractions (one case)."""
    return l1_errors(case, a32, a64)


def _pool_all(cases, precision, args):
    """All cases admitted to one pool in ``precision`` ("fp32s", "fp64" or
    "routed": each case as route_precision says) and run to their t_final.
    Returns (results, per-case finish times, wall seconds, pool stats)."""
    from ecogen_amd.pool import Pool, route_precision
    finish = {}
    with Pool(hbm_budget=96 << 30, steps_per_chunk=24, use_graph=not args.batch_direct) as pool:
        t0 = time.perf_counter()
        ids = []
        for c, st, tf in cases:
            prec = route_precision(c, st, tf) if precision == "routed" else precision
            ids.append(pool.admit(dict(c, precision=prec), st, t_final=tf))
        if any(i is None for i in ids):
            raise SystemExit("batched leg: cases exceed the pool budget")
        while pool.stats().running > 0:
            pool.advance()
            for cid in pool.poll(keep=1):
                finish[cid] = time.perf_counter() - t0
        pool.poll(keep=0)
        wall = time.perf_counter() - t0
        for cid in pool.finished:
            finish.setdefault(cid, wall)
        stats = pool.stats()
        res = [pool.result(i) for i in ids]
    return res, [finish[i] for i in ids], wall, stats


def leg_batched(args):
    """2048 independent water/air Riemann problems (70 % 1-D, 30 % 2-D), Kapila
    MUSCL van Leer, each case run to its own t_final in one pool by persistent
    batched kernels replayed as hipGraph chunks: all in FP32 storage (FP64
    relaxation), all in FP64 (the like-for-like comparator), and routed (FP32
    storage only where it stays within 1e-5 of FP64); the FP32-storage error of
    every result against its FP64 result; and a sample solved one at a time
    with the plain FP64 path."""
    from ecogen_amd import Context
    from ecogen_amd import workloads as W
    from ecogen_amd.pool import route_precision, FP32S_MAX_STEPS_1D
    specs = W.batch_specs(args.batch_cases)
    cases = [W.riemann_case(*sp, precision="fp32s") for sp in specs]
    cells = np.array([int(np.prod(c["cells"])) for c, _, _ in cases])
    routes = [route_precision(c, st, tf) for c, st, tf in cases]

    def summary(res, lat, wall, stats):
        steps = np.array([r[1].steps for r in res])
        cu = float((cells * steps).sum())
        return {"wall_s": round(wall, 3), "solves_per_s": round(len(cases) / wall, 2),
                "cell_updates_per_s": round(cu / wall, 1), "cell_updates": cu,
                "latency_s": {"p50": round(_pct(lat, 50), 3), "p99": round(_pct(lat, 99), 3), "max": round(max(lat), 3)},
                "relax_not_converged_cases": int(sum(1 for r in res if r[1].flags & 1)),
                "graph_captures": stats.graph_captures, "tiles_launched": stats.tiles_launched}

    r32, lat32, wall32, st32 = _pool_all(cases, "fp32s", args)
    s32 = summary(r32, lat32, wall32, st32)
    log(f"batched fp32s: {wall32:.1f} s")
    r64, lat64, wall64, st64 = _pool_all(cases, "fp64", args)
    s64 = summary(r64, lat64, wall64, st64)
    log(f"batched fp64: {wall64:.1f} s")
    rr, latr, wallr, str_ = _pool_all(cases, "routed", args)
    sr = summary(rr, latr, wallr, str_)
    log(f"batched routed: {wallr:.1f} s")
    # FP32-storage error of every case against its FP64 result (same step count required)
    l1 = []
    for k, (c, st, tf) in enumerate(cases):
        if r32[k][1].steps == r64[k][1].steps:
            e = _fp32s_vs_fp64(c, r32[k][0], r64[k][0])
            l1.append((k, max(e.values()), bool((r32[k][1].flags | r64[k][1].flags) & 1)))
        else:
            l1.append((k, float("inf"), bool((r32[k][1].flags | r64[k][1].flags) & 1)))
    worst = np.array([x[1] for x in l1])
    relax = np.array([x[2] for x in l1])
    in32 = np.array([rt == "fp32s" for rt in routes])
    # routed results against FP64: the FP64-routed ones equal it bit for bit
    routed_fp64_equal = all(np.array_equal(rr[k][0], r64[k][0]) for k in range(len(cases)) if not in32[k])
    routed_l1 = [float(max(_fp32s_vs_fp64(c, rr[k][0], r64[k][0]).values())) for k, (c, _, _) in enumerate(cases) if in32[k]]
    # serial plain FP64 solves of a sample (create, upload, run to t_final, download, destroy)
    sample = [] if args.no_serial else list(range(0, len(cases), max(1, args.serial_stride)))
    ser_t, ser_cu = 0.0, 0.0
    for k in sample:
        c, st, tf = cases[k]
        t1 = time.perf_counter()
        with Context(dict(c, precision="fp64")) as ctx:
            ctx.set_state(st)
            s64k = ctx.run(t_final=tf)
            ctx.get_state()
        ser_t += time.perf_counter() - t1
        ser_cu += cells[k] * s64k.steps
    ser_solves = len(sample) / ser_t if sample else None
    ser_cus = ser_cu / ser_t if sample else None
    return {
        "cases": len(cases), "cells": int(cells.sum()),
        "fp32s": s32, "fp64": s64, "routed": sr,
        "fp32s_speedup_vs_fp64_pool": {"cell_updates": round(s32["cell_updates_per_s"] / s64["cell_updates_per_s"], 3),
                                       "solves": round(s32["solves_per_s"] / s64["solves_per_s"], 3)},
        "routed_speedup_vs_fp64_pool": round(wall64 / wallr, 3),
        "fp32s_l1_vs_fp64_every_case": {
            "median": float(f"{np.median(worst[np.isfinite(worst)]):.3e}"),
            "within_1e-5": int((worst <= 1e-5).sum()), "over_1e-5": int((worst > 1e-5).sum()),
            "over_1e-5_without_relaxation_failures": int(((worst > 1e-5) & ~relax).sum()),
            "relaxation_failure_cases": int(relax.sum()),
            "quantity": "max over mixture density, mixture pressure, velocity, volume fractions"},
        "routing": {"rule": f"FP32 storage for 1-D two-phase MUSCL transmissive runs of at most {FP32S_MAX_STEPS_1D} "
                            "estimated steps (t_final over the first CFL step); FP64 otherwise",
                    "fp32s_cases": int(in32.sum()), "fp64_cases": int((~in32).sum()),
                    "fp32s_l1_max": float(f"{max(routed_l1):.3e}") if routed_l1 else None,
                    "fp32s_within_1e-5": int(sum(1 for x in routed_l1 if x <= 1e-5)),
                    "fp64_routed_equal_to_fp64_pool": bool(routed_fp64_equal)},
        "serial_fp64": {"sample": f"every {args.serial_stride}th case ({len(sample)} cases), one solve at a time",
                        "solves_per_s": round(ser_solves, 2) if sample else None,
                        "cell_updates_per_s": round(ser_cus, 1) if sample else None},
        "fp64_pool_speedup_vs_serial_fp64": {"solves": round(s64["solves_per_s"] / ser_solves, 2) if sample else None,
                                             "cell_updates": round(s64["cell_updates_per_s"] / ser_cus, 2) if sample else None},
        "storage": "FP32 storage: alpha pair (FP64 in two words), rho_k, p_k, u, reduced energy in FP32; FP64 fluxes, "
                   "update, relaxation",
    }


def leg_served(args):
    """Poisson request streams (same requests at every rate) through the C++
    serving loop: continuous batching, admission by a 200 GB HBM budget;
    against a serial plain FP64 server (measured service time of every request,
    first-in first-out)."""
    from ecogen_amd import Context
    from ecogen_amd import workloads as W
    from ecogen_amd.pool import Pool
    from ecogen_amd.pool import route_precision
    base = W.request_stream(1.0, args.serve_requests)
    reqcases = [W.request_case(kind, prm, precision="fp32s") for _, kind, prm in base]
    # each request in the precision the routing rule gives it (FP32 storage
    # only where it stays within 1e-5 of FP64)
    reqcases = [(dict(c, precision=route_precision(c, st, tf)), st, tf) for c, st, tf in reqcases]
    cells = np.array([int(np.prod(c["cells"])) for c, _, _ in reqcases])
    # serial plain FP64 service times.  Each solve is capped at four times its
    # CFL step estimate (t_final = 0.3 L / a_max: about 0.3 n / (cfl * lcfl
    # factor) steps) plus 200: a request whose FP64 trajectory collapses dt (a
    # chaotic 3-D bubble collapse can) then counts with a lower bound of its
    # service time, which favours the serial comparator.
    service, capped = [], 0
    for c, st, tf in reqcases:
        dim = sum(1 for n in c["cells"] if n > 1)
        est = 0.3 * max(c["cells"]) / (c["cfl"] * (0.6 if dim > 1 else 1.0))
        t1 = time.perf_counter()
        with Context(dict(c, precision="fp64")) as ctx:
            ctx.set_state(st)
            s64 = ctx.run(t_final=tf, max_steps=int(4 * est) + 200)
            ctx.get_state()
        service.append(time.perf_counter() - t1)
        capped += int(s64.t < float(np.float32(tf)))
    out = {"requests_per_rate": len(reqcases), "bubble_3d_requests": sum(1 for _, k, _ in base if k == "bubbles"),
           "fp32s_requests": sum(1 for c, _, _ in reqcases if c["precision"] == "fp32s"),
           "hbm_budget_gb": 200, "serial_fp64_capped_requests": capped,
           "serial_fp64_step_cap": "4x the CFL step estimate + 200",
           "rates": {}}
    for lam in args.serve_rates:
        arrivals = [a / lam for a, _, _ in base]
        reqs = [{"case": c, "state": st, "t_final": tf, "arrival_s": arr}
                for (c, st, tf), arr in zip(reqcases, arrivals)]
        with Pool(hbm_budget=200 << 30, steps_per_chunk=24) as pool:
            t0 = time.perf_counter()
            res, _ = pool.serve(reqs)
            wall = time.perf_counter() - t0
        tts = res[:, 1] - np.array(arrivals)
        cu = float((cells * res[:, 2]).sum())
        # FIFO serial FP64 server on the same arrivals
        fin, ser = 0.0, []
        for arr, sv in zip(arrivals, service):
            fin = max(fin, arr) + sv
            ser.append(fin - arr)
        out["rates"][str(int(lam))] = {
            "tts_s": {"p50": round(_pct(tts, 50), 3), "p99": round(_pct(tts, 99), 3)},
            "solves_per_s": round(len(reqs) / wall, 2), "cell_updates_per_s": round(cu / wall, 1),
            "wall_s": round(wall, 3),
            "serial_fp64_tts_s": {"p50": round(_pct(ser, 50), 3), "p99": round(_pct(ser, 99), 3)},
            "serial_fp64_solves_per_s": round(len(reqs) / fin, 2),
        }
    return out


def leg_latency_1d():
    """1-D Kapila first-order water/air shock tube, 1000 cells, FP64, to 240 us:
    time to solution of one solve (upload, run, download) and per step."""
    from ecogen_amd import Context
    from ecogen_amd import workloads as W
    case, st = W.shock_tube_1d()
    with Context(case) as ctx:  # warm-up (module load, first launches)
        ctx.set_state(st)
        ctx.run(t_final=case["t_final"])
    t0 = time.perf_counter()
    with Context(case) as ctx:
        ctx.set_state(st)
        s = ctx.run(t_final=case["t_final"])
        ctx.get_state()
    wall = time.perf_counter() - t0
    return {"workload": "1-D Kapila FO shock tube, 1000 cells, FP64, t=240 us", "steps": s.steps,
            "time_to_solution_s": round(wall, 4), "us_per_step": round(wall / s.steps * 1e6, 2),
            "cell_updates_per_s": round(1000 * s.steps / wall, 1)}


def leg_c3(args):
    """3-D pressure-disequilibrium model with stiff relaxation, MUSCL minmod,
    64 air bubbles in water at 5e7 Pa: FP64 and FP32 storage."""
    from ecogen_amd import workloads as W
    n = args.c3_n
    s = n / 384.0
    case, st = W.bubbles_3d(n=n, rmin=4.0 * s, rmax=24.0 * s)
    sub = argparse.Namespace(warmup=2, steps=args.c3_steps)
    r64 = run_precision(case, st, sub, "fp64")
    r32 = run_precision(case, st, sub, "fp32s")
    del st
    ncell = n ** 3
    return {"workload": f"3-D MultiP + relaxation, MUSCL minmod, {n}^3, 64 bubbles", "steps": args.c3_steps,
            "fp64": {"cell_updates_per_s": round(r64["value"], 1), "ms_per_step": round(r64["wall"] / args.c3_steps * 1e3, 3),
                     "roofline": roofline(case, r64, 8)},
            "fp32s": {"cell_updates_per_s": round(r32["value"], 1), "ms_per_step": round(r32["wall"] / args.c3_steps * 1e3, 3),
                      "roofline": roofline(case, r32, 4)},
            "fp32s_speedup": round(r32["value"] / r64["value"], 3),
            "fp32s_l1_vs_fp64": {k: float(f"{v:.3e}") for k, v in l1_errors(case, r32["state"], r64["state"]).items()},
            "cells": ncell}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the step runs on one GPU: sharding across GPUs (slabs of the grid "
                                   "with a two-cell halo exchange per stage) is not built. Use --gpus 1."}))
        sys.exit(2)
    from ecogen_amd import workloads as W

    log(f"headline workload {args.n}^2, {args.steps} steps")
    case, st = W.shock_cylinders_2d(n=args.n, steps=args.steps)
    ncell = int(np.prod(case["cells"]))
    r64 = run_precision(case, st, args, "fp64")
    log(f"fp64: {r64['wall'] / args.steps * 1e3:.3f} ms/step")
    roof = roofline(case, r64, 8)
    roof["traffic"] = traffic_from_profiles(args.n, "fp64", roof["kernel"])
    roof["valu"] = valu_from_profiles(args.n, "fp64", roof["kernel"], roof["kernel_ms"])
    fp32s = None
    if not args.no_fp32s:
        r32 = run_precision(case, st, args, "fp32s")
        roof32 = roofline(case, r32, 4)
        roof32["traffic"] = traffic_from_profiles(args.n, "fp32s", roof32["kernel"])
        roof32["valu"] = valu_from_profiles(args.n, "fp32s", roof32["kernel"], roof32["kernel_ms"])
        fp32s = {"value": round(r32["value"], 1), "speedup_vs_fp64": round(r32["value"] / r64["value"], 3),
                 "ms_per_step": round(r32["wall"] / args.steps * 1e3, 4),
                 "l1_vs_fp64": {k: float(f"{v:.3e}") for k, v in l1_errors(case, r32["state"], r64["state"]).items()},
                 "roofline": roof32,
                 "storage": "alpha_k, rho_k, p_k, u, reduced energy in FP32; fluxes, update, relaxation in FP64"}
    del st
    if args.calibrate_copy:
        copy_calibration()
    extra = {}
    legs = set() if args.only_main else set(args.legs)
    if legs:
        if "latency" in legs:
            log("1-D shock-tube latency")
            extra["latency_1d"] = leg_latency_1d()
        if "batched" in legs and args.batch_cases > 0:
            log(f"batched pool: {args.batch_cases} cases")
            extra["batched"] = leg_batched(args)
            log(f"batched: fp32s {extra['batched']['fp32s']['wall_s']} s, fp64 {extra['batched']['fp64']['wall_s']} s")
        if "served" in legs and args.serve_requests > 0:
            log(f"served stream: {args.serve_requests} requests x {len(args.serve_rates)} rates")
            extra["served"] = leg_served(args)
        if "c3" in legs and args.c3_steps > 0:
            log(f"3-D bubbles {args.c3_n}^3, {args.c3_steps} steps")
            extra["c3"] = leg_c3(args)
    if not args.no_cpu:
        log("CPU baseline")
    cpu = None if args.no_cpu else cpu_baseline(args)
    line = {
        "metric": METRIC, "value": round(r64["value"], 1), "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(r64["wall"] / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2-D Kapila 5-eq MUSCL vanLeer FP64 {args.n}x{args.n}, Mach 2.4 air shock on 32 "
                               f"water cylinders", "cells": ncell, "model": "kapila", "order": 2,
                   "limiter": "vanleer", "phases": 2, "parallelism": "single GPU",
                   "t_end": r64["status"].t},
        "roofline": roof,
        "cpu_baseline": cpu,
        "fp32_storage": fp32s,
    }
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
