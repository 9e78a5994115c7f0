"""Tile-shape and occupancy autotuner for the stage kernels.

Three steps, so that the compile happens on the development host and the
timing on the GPU:

    python -m ecogen_amd.autotune build            # compile every variant (hipcc, gfx950)
    python -m ecogen_amd.autotune run  --out R.json # on the GPU: time each variant
    python -m ecogen_amd.autotune apply R.json      # write the fastest into csrc/eco_tuning.h

A variant is a set of build-time macros of ``csrc/eco_tuning.h`` (tile
shape and workgroup size per dimension, minimum waves per SIMD requested
from the compiler for the single-solve and the pool kernels).
Each is timed on the stage kernels of a fixed workload with HIP events
(eco_time_steps), or, for the pool kernels, by the wall time of a pool
running the batched Riemann workload; its results must equal the default
build's on the same states.
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
TUNE_DIR = PKG / "_tune"
CSRC = PKG / "csrc"

# (tile cells per dimension, workgroup threads): a workgroup may be larger
# than its tile so that the face pass ends in fewer iterations
SPACE = {
    1: {"shapes": [((504,), 512), ((248,), 256), ((1016,), 1024), ((256,), 256)], "waves": [1, 2]},
    3: {"shapes": [((16, 4, 3), 256), ((16, 4, 4), 256), ((8, 8, 4), 256)], "waves": [2, 3]},
}
# 2-D: the staged MUSCL pass runs one face per thread and direction, so a
# tile's faces must fit its workgroup; (tile, threads, waves per SIMD of the
# staged kernels).  The shapes keep the faces of each direction, the cells'
# limited slopes and the cells themselves close to whole waves.
# (tile, threads, waves per SIMD of the FP64 and of the FP32-storage staged kernels; the
# FP32-storage staged kernel keeps its own tile height, ECO_TILE32S_2_Y)
SPACE2 = [((16, 12), 256, 4, 4), ((16, 13), 256, 4, 4), ((16, 13), 256, 3, 4), ((16, 14), 256, 3, 4),
          ((16, 15), 256, 3, 4)]
AXES = "XYZ"
# "dimension" 0: the persistent pool kernels of the FP32-storage form, timed
# on the batched Riemann workload (minimum waves per SIMD of the 1-D and
# 2-D pool kernels; 0 keeps the single-solve request)
POOL_SPACE = {"b1": [0, 4], "b2": [0, 3]}


def variants(dim):
    if dim == 0:
        return [(f"pool_b1{w1}_b2{w2}", {"ECO_WAVESB32_1": w1, "ECO_WAVESB32_2": w2})
                for w1, w2 in itertools.product(POOL_SPACE["b1"], POOL_SPACE["b2"])]
    out = []
    if dim == 2:
        for t, nt, w, w32 in SPACE2:
            name = f"d2_t{t[0]}x{t[1]}_n{nt}_ws{w}_{w32}"
            out.append((name, {"ECO_TILE2_X": t[0], "ECO_TILE2_Y": t[1], "ECO_THREADS2": nt, "ECO_WAVESS2": w,
                               "ECO_WAVESS32_2": w32, "ECO_LIMSPEC2": 1}))
        return out
    sp = SPACE[dim]
    for (t, nt), w in itertools.product(sp["shapes"], sp["waves"]):
        name = f"d{dim}_t{'x'.join(map(str, t))}_n{nt}_w{w}"
        defs = {f"ECO_TILE{dim}_{AXES[a]}": t[a] for a in range(dim)}
        defs.update({f"ECO_THREADS{dim}": nt, f"ECO_WAVES{dim}": w, f"ECO_LIMSPEC{dim}": 1})
        out.append((name, defs))
    return out


# the translation unit whose kernels a dimension's workload runs; the other
# objects of a variant are those of the main build (python -m ecogen_amd.build)
TU_FOR_DIM = {0: ("eco_k_kapila32.hip",), 1: ("eco_k_kapila32.hip",), 2: ("eco_k_kapila.hip", "eco_k_kapila32.hip"),
              3: ("eco_k_multip.hip",)}


def _compile_variant(item):
    from .build import ARCH, COMMON, HIPCC, OBJ, SOURCES
    dim, name, defs = item
    vdir = TUNE_DIR / name
    vdir.mkdir(parents=True, exist_ok=True)
    dflags = [f"-D{k}={v}" for k, v in defs.items()]
    objs, fresh = [], []
    for src in SOURCES:
        if src not in (*TU_FOR_DIM[dim], "eco_dispatch.cpp"):
            objs.append(str(OBJ / (Path(src).stem + ".o")))
            continue
        obj = vdir / (Path(src).stem + ".o")
        cmd = [HIPCC, *COMMON, *dflags, "-c", str(CSRC / src), "-o", str(obj)]
        if src.endswith(".cpp"):
            cmd[1:1] = ["-x", "hip"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            return name, f"compile failed: {r.stderr[-800:]}"
        objs.append(str(obj))
        fresh.append(obj)
    lib = vdir / "libecogen_amd.so"
    r = subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", str(lib)],
                       capture_output=True, text=True)
    for o in fresh:
        os.remove(o)
    return name, ("ok" if r.returncode == 0 else f"link failed: {r.stderr[-800:]}")


def cmd_build(args):
    items = [(d, name, defs) for d in args.dims for name, defs in variants(d)]
    with ThreadPoolExecutor(max_workers=args.jobs) as ex:
        for name, status in ex.map(_compile_variant, items):
            print(f"{name}: {status}", flush=True)


def _workload(dim, n):
    from . import workloads as W
    if dim == 1:
        case, st, _ = W.riemann_case(1, n, 100.0, 0.5, precision="fp32s")
        return case, st
    if dim == 2:
        s = n / 4096.0
        return W.shock_cylinders_2d(n=n, rmin=20.0 * s, rmax=80.0 * s)
    s = n / 384.0
    return W.bubbles_3d(n=n, rmin=max(1.5, 4.0 * s), rmax=24.0 * s)


def _time_lib(lib_path, case, st, steps):
    """Load one variant library in a fresh namespace and time its stage kernels."""
    from . import capi
    capi._lib = None
    capi.LIB_PATH = Path(lib_path)
    ctx = capi.Context(case)
    ctx.set_state(st)
    ctx.run(nsteps=2)
    status, ms, cnt = ctx.time_steps(steps)
    out = ctx.get_state()
    ctx.close()
    capi._lib = None
    return float(ms.sum()) / steps, out


def _time_pool(lib_path, count):
    """Load one variant library and run ``count`` cases of the batched
    Riemann workload through a pool: (wall seconds, per-case results)."""
    import time
    from . import capi
    from . import pool as P
    from . import workloads as W
    capi._lib, P._bound = None, False
    capi.LIB_PATH = Path(lib_path)
    cases = [W.riemann_case(*sp, precision="fp32s") for sp in W.batch_specs(count)]
    with P.Pool(hbm_budget=48 << 30, steps_per_chunk=24) as pool:
        ids = [pool.admit(c, st, t_final=tf) for c, st, tf in cases]
        t0 = time.perf_counter()
        pool.run_all()
        wall = time.perf_counter() - t0
        outs = [pool.result(i)[0] for i in ids]
    capi._lib, P._bound = None, False
    return wall, outs


def _run_pool(args, results):
    base_s, base = _time_pool(PKG / "libecogen_amd.so", args.pool_cases)
    print(f"pool default build: {base_s:.2f} s for {args.pool_cases} cases", flush=True)
    for name, defs in variants(0):
        lib = TUNE_DIR / name / "libecogen_amd.so"
        if not lib.exists():
            continue
        wall, outs = _time_pool(lib, args.pool_cases)
        dev = max(float(np.abs(a - b).max()) for a, b in zip(outs, base))
        results.append({"dim": 0, "variant": name, "defs": defs, "wall_ms": wall * 1e3, "max_abs_diff": dev})
        print(f"  {name}: {wall:.2f} s  (max |diff| vs default {dev:.2e})", flush=True)


def cmd_run(args):
    results = []
    for dim in args.dims:
        if dim == 0:
            _run_pool(args, results)
            continue
        case, st = _workload(dim, {1: args.n1, 2: args.n2, 3: args.n3}[dim])
        # 2-D: the headline workload in FP64 and in FP32 storage (one tile shape for both)
        forms = [(case, "")] + ([(dict(case, precision="fp32s"), " fp32s")] if dim == 2 else [])
        for cs, tag in forms:
            base_ms, base = _time_lib(PKG / "libecogen_amd.so", cs, st, args.steps)
            print(f"dim {dim}{tag} default build: {base_ms:.3f} ms/step", flush=True)
            for name, defs in variants(dim):
                lib = TUNE_DIR / name / "libecogen_amd.so"
                if not lib.exists():
                    continue
                ms, out = _time_lib(lib, cs, st, args.steps)
                dev = float(np.abs(out - base).max())
                results.append({"dim": dim, "form": tag.strip() or "fp64", "variant": name, "defs": defs,
                                "ms_per_step": ms, "max_abs_diff": dev})
                print(f"  {name}{tag}: {ms:.3f} ms/step  (max |diff| vs default {dev:.2e})", flush=True)
    Path(args.out).parent.mkdir(parents=True, exist_ok=True)
    Path(args.out).write_text(json.dumps(results, indent=1))


def _cost(r):
    return r["ms_per_step"] if "ms_per_step" in r else r["wall_ms"]


def cmd_apply(args):
    res = json.loads(Path(args.results).read_text())
    best = {}
    for r in res:
        if r["max_abs_diff"] != 0.0:
            continue  # a tile shape must not change the result
        if r["dim"] not in best or _cost(r) < _cost(best[r["dim"]]):
            best[r["dim"]] = r
    hdr = (CSRC / "eco_tuning.h").read_text()
    for dim, r in sorted(best.items()):
        for k, v in r["defs"].items():
            hdr = re.sub(rf"#define {k} \w+", f"#define {k} {v}", hdr)
        print(f"dim {dim}: {r['variant']} {_cost(r):.3f} ms")
    (CSRC / "eco_tuning.h").write_text(hdr)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    b = sub.add_parser("build")
    b.add_argument("--dims", type=int, nargs="+", default=[0, 1, 2, 3])
    b.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    r = sub.add_parser("run")
    r.add_argument("--dims", type=int, nargs="+", default=[0, 1, 2, 3])
    r.add_argument("--n1", type=int, default=1 << 22)
    r.add_argument("--n2", type=int, default=4096)
    r.add_argument("--n3", type=int, default=256)
    r.add_argument("--steps", type=int, default=10)
    r.add_argument("--pool-cases", type=int, default=512)
    r.add_argument("--out", default="autotune.json")
    a = sub.add_parser("apply")
    a.add_argument("results")
    args = ap.parse_args(argv)
    {"build": cmd_build, "run": cmd_run, "apply": cmd_apply}[args.cmd](args)


if __name__ == "__main__":
    main()
