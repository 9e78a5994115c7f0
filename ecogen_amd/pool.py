"""ctypes binding of the pool / serving C ABI (include/ecogen_pool.h).

A :class:`Pool` keeps many solves resident on the GPU and advances them
together (persistent batched kernels, hipGraph-replayed chunks of steps);
:meth:`Pool.serve` runs a timed request stream through the C++ serving loop.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .capi import EcoConfig, EcoError, EcoStatus, lib, make_config, nfields

ERR_BUDGET = -6
ERR_FULL = -7

POOL_EXPORTS = ["eco_pool_create", "eco_pool_case_bytes", "eco_pool_admit", "eco_pool_advance", "eco_pool_poll",
                "eco_pool_result", "eco_pool_release", "eco_pool_stats_get", "eco_pool_destroy", "eco_serve"]


class PoolConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_ulonglong), ("max_cases", C.c_int), ("steps_per_chunk", C.c_int),
                ("use_graph", C.c_int), ("grid_blocks", C.c_int)]


class PoolStats(C.Structure):
    _fields_ = [("active", C.c_int), ("running", C.c_int), ("bytes_used", C.c_ulonglong),
                ("bytes_free", C.c_ulonglong), ("steps_launched", C.c_longlong), ("tiles_launched", C.c_longlong),
                ("graph_captures", C.c_int)]


class Request(C.Structure):
    _fields_ = [("cfg", C.POINTER(EcoConfig)), ("state", C.POINTER(C.c_double)), ("nsteps", C.c_int),
                ("t_final", C.c_double), ("arrival_s", C.c_double), ("out_state", C.POINTER(C.c_double))]


class RequestResult(C.Structure):
    _fields_ = [("admit_s", C.c_double), ("finish_s", C.c_double), ("steps", C.c_int), ("t", C.c_double),
                ("flags", C.c_uint32)]


_bound = False


def _bind():
    global _bound
    L = lib()
    if _bound:
        return L
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    L.eco_pool_create.argtypes = [C.POINTER(PoolConfig), C.POINTER(vp)]
    L.eco_pool_create.restype = C.c_int
    L.eco_pool_case_bytes.argtypes = [C.POINTER(EcoConfig)]
    L.eco_pool_case_bytes.restype = C.c_ulonglong
    L.eco_pool_admit.argtypes = [vp, C.POINTER(EcoConfig), dp, C.c_int, C.c_double, C.POINTER(C.c_int)]
    L.eco_pool_admit.restype = C.c_int
    L.eco_pool_advance.argtypes = [vp]
    L.eco_pool_advance.restype = C.c_int
    L.eco_pool_poll.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]
    L.eco_pool_poll.restype = C.c_int
    L.eco_pool_result.argtypes = [vp, C.c_int, dp, C.POINTER(EcoStatus)]
    L.eco_pool_result.restype = C.c_int
    L.eco_pool_release.argtypes = [vp, C.c_int]
    L.eco_pool_release.restype = C.c_int
    L.eco_pool_stats_get.argtypes = [vp, C.POINTER(PoolStats)]
    L.eco_pool_stats_get.restype = C.c_int
    L.eco_pool_destroy.argtypes = [vp]
    L.eco_pool_destroy.restype = None
    L.eco_serve.argtypes = [vp, C.POINTER(Request), C.c_int, C.POINTER(RequestResult)]
    L.eco_serve.restype = C.c_int
    _bound = True
    return L


def _check(rc, where):
    if rc != 0:
        raise EcoError(rc, where)


def _host_state(case, state):
    nf = nfields(case)
    ncell = int(np.prod(case["cells"]))
    s = np.zeros((nf, ncell))
    st = np.asarray(state, dtype=np.float64)
    s[: st.shape[0]] = st.reshape(st.shape[0], ncell)
    return np.ascontiguousarray(s)


class Pool:
    """Many resident solves advanced together on one GPU."""

    def __init__(self, hbm_budget=8 << 30, max_cases=8192, steps_per_chunk=24, use_graph=True, grid_blocks=0):
        L = _bind()
        self.pc = PoolConfig(int(hbm_budget), int(max_cases), int(steps_per_chunk), int(bool(use_graph)),
                             int(grid_blocks))
        h = C.c_void_p()
        _check(L.eco_pool_create(C.byref(self.pc), C.byref(h)), "eco_pool_create")
        self._h = h
        self.cases = {}
        self.finished = []  # handles reported finished, in order

    def close(self):
        if getattr(self, "_h", None):
            lib().eco_pool_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @staticmethod
    def case_bytes(case):
        cfg = make_config(case)
        return int(_bind().eco_pool_case_bytes(C.byref(cfg)))

    def admit(self, case, state, nsteps=0, t_final=0.0):
        """Admit one solve; returns its handle, or None when it does not fit the budget now."""
        cfg = make_config(case)
        s = _host_state(case, state)
        cid = C.c_int(-1)
        rc = _bind().eco_pool_admit(self._h, C.byref(cfg), s.ctypes.data_as(C.POINTER(C.c_double)), int(nsteps),
                                    float(t_final), C.byref(cid))
        if rc in (ERR_BUDGET, ERR_FULL):
            return None
        _check(rc, "eco_pool_admit")
        self.cases[cid.value] = case
        return cid.value

    def advance(self):
        _check(lib().eco_pool_advance(self._h), "eco_pool_advance")

    def poll(self, keep=0, max_ids=4096):
        ids = (C.c_int * max_ids)()
        n = C.c_int(0)
        _check(lib().eco_pool_poll(self._h, int(keep), ids, max_ids, C.byref(n)), "eco_pool_poll")
        out = [ids[i] for i in range(n.value)]
        self.finished += out
        return out

    def result(self, cid):
        case = self.cases[cid]
        s = np.zeros((nfields(case), int(np.prod(case["cells"]))))
        st = EcoStatus()
        _check(lib().eco_pool_result(self._h, int(cid), s.ctypes.data_as(C.POINTER(C.c_double)), C.byref(st)),
               "eco_pool_result")
        return s, st

    def release(self, cid):
        _check(lib().eco_pool_release(self._h, int(cid)), "eco_pool_release")
        self.cases.pop(cid, None)
        if cid in self.finished:
            self.finished.remove(cid)

    def stats(self):
        s = PoolStats()
        _check(lib().eco_pool_stats_get(self._h, C.byref(s)), "eco_pool_stats_get")
        return s

    def run_all(self, keep=1):
        """Advance until no admitted case is running; returns every handle reported
        finished so far, in finishing order."""
        while self.stats().running > 0:
            self.advance()
            self.poll(keep=keep)
        self.poll(keep=0)
        return list(self.finished)

    def serve(self, requests, keep_states=False):
        """Run a request stream (list of dicts: case, state, t_final, nsteps, arrival_s) through the
        C++ serving loop in real time.  Returns (results array of (admit_s, finish_s, steps, t, flags),
        list of final states or None)."""
        n = len(requests)
        reqs = (Request * n)()
        keep = []  # keep ctypes objects alive
        outs = []
        for i, r in enumerate(requests):
            cfg = make_config(r["case"])
            s = _host_state(r["case"], r["state"])
            out = np.zeros_like(s) if keep_states else None
            keep += [cfg, s, out]
            outs.append(out)
            reqs[i].cfg = C.pointer(cfg)
            reqs[i].state = s.ctypes.data_as(C.POINTER(C.c_double))
            reqs[i].nsteps = int(r.get("nsteps", 0))
            reqs[i].t_final = float(r.get("t_final", 0.0))
            reqs[i].arrival_s = float(r["arrival_s"])
            reqs[i].out_state = out.ctypes.data_as(C.POINTER(C.c_double)) if out is not None else None
        res = (RequestResult * n)()
        _check(lib().eco_serve(self._h, reqs, n, res), "eco_serve")
        arr = np.array([(x.admit_s, x.finish_s, x.steps, x.t, x.flags) for x in res])
        return arr, outs


# Precision routing of pooled and served solves.  FP32 storage is held to
# 1e-5 relative L1 of the FP64 result.  On the batched water/air workload
# (2048 cases, every result compared with its FP64 run) its error grows with
# the step count -- median 2e-7 below 200 steps, 8e-6 at 2000-5000 steps,
# 3e-5 beyond -- and 2-D columns whose expansion puts the water in tension
# (the relaxation then stops at its iteration limit in both precisions) end
# up to 0.8 apart.  Only 1-D solves of at most FP32S_MAX_STEPS_1D estimated
# steps stayed within 1e-5 in every case there; everything else runs in FP64.
FP32S_MAX_STEPS_1D = 100


def estimated_steps(case, state, t_final):
    """t_final over the first CFL step cfl * lcfl / a_max of the initial state
    (a_max: the largest |u| + frozen sound speed of a cell; lcfl: the
    smallest cell size, x 0.6 in multi-D, as the step uses)."""
    if not t_final or t_final <= 0:
        return float(case.get("steps", 0))
    st = np.asarray(state, dtype=np.float64)
    eos = case["eos"]
    nph = len(eos)
    if case["model"] == "euler":
        rho = st[1]
        c2 = eos[0]["gamma"] * (st[2] + eos[0].get("pinf", 0.0)) / rho
        u = np.abs(st[3:6]).max(axis=0)
    else:
        rho = sum(st[3 * k] * st[3 * k + 1] for k in range(nph))
        c2 = sum(st[3 * k] * st[3 * k + 1] / rho * eos[k]["gamma"] * (st[3 * k + 2] + eos[k].get("pinf", 0.0))
                 / st[3 * k + 1] for k in range(nph))
        u = np.abs(st[3 * nph:3 * nph + 3]).max(axis=0)
    a_max = float((u + np.sqrt(np.maximum(c2, 0.0))).max())
    dims = [d for d in range(3) if case["cells"][d] > 1]
    lcfl = min(case["lengths"][d] / case["cells"][d] for d in dims) * (0.6 if len(dims) > 1 else 1.0)
    return t_final / (case["cfl"] * lcfl / a_max) if a_max > 0 else float("inf")


def route_precision(case, state, t_final=0.0):
    """"fp32s" for a solve FP32 storage keeps within 1e-5 of FP64 (a 1-D
    two-phase MUSCL transmissive run of at most FP32S_MAX_STEPS_1D estimated
    steps), else "fp64"."""
    dim = sum(1 for n in case["cells"] if n > 1)
    fp32_form = (case["model"] != "euler" and case["order"] == 2 and len(case["eos"]) == 2
                 and set(case.get("bc", ["abs"])) <= {"abs"})
    if fp32_form and dim == 1 and estimated_steps(case, state, t_final) <= FP32S_MAX_STEPS_1D:
        return "fp32s"
    return "fp64"
