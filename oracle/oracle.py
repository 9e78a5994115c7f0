"""ctypes front end of the CPU oracle (test infrastructure only).

The oracle is a plain-C FP64 restatement of the reference ECOGEN step
(``oracle/eco_oracle.c``).  Tests, the smoke check and the CPU-baseline leg of
the benchmark load it here, as the checker.  Nothing in the product imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libeco_oracle.so"

MODELS = {"euler": 0, "kapila": 1, "multip": 2, "thermaleq": 3, "eulerhomogeneous": 4}
# models whose host state ends with the mixture temperature
WITH_T = ("thermaleq", "eulerhomogeneous")
LIMITERS = {"none": 0, "minmod": 1, "vanleer": 2, "vanalbada": 3, "superbee": 4, "mc": 5, "thinc": 6}
EOS_TYPES = {"IG": 0, "SG": 1}
# per side X-, X+, Y-, Y+, Z-, Z+; outflow pressures in case["bc_p0"]; injections in
# case["bc_inj"][side] = {"m0": mass flow into the domain, "alpha", "rho", "p": per phase}
# case["bc_tank"][side] = {"p0", "T0", and per phase "alpha" or "massFrac"}
BCS = {"abs": 0, "wall": 1, "outflow": 2, "injection": 3, "tank": 4}
MAXPH = 4
# transported scalars: case["transports"] = names, each state's "tr" = values
# in that order, case["bc_tr"][side] = values at an outflow, injection or tank side
MAXTR = 4


class OrcEos(C.Structure):
    _fields_ = [("type", C.c_int), ("gamma", C.c_double), ("pinf", C.c_double),
                ("cv", C.c_double), ("eref", C.c_double)]


class OrcCfg(C.Structure):
    _fields_ = [("model", C.c_int), ("nph", C.c_int), ("order", C.c_int), ("lim", C.c_int * 4),
                ("alpha_null", C.c_int), ("relax_p", C.c_int), ("n", C.c_int * 3),
                ("L", C.c_double * 3), ("cfl", C.c_double), ("eos", OrcEos * MAXPH), ("bc", C.c_int * 6),
                ("bc_p0", C.c_double * 6), ("bc_m0", C.c_double * 6),
                ("bc_ak0", (C.c_double * MAXPH) * 6), ("bc_rhok0", (C.c_double * MAXPH) * 6),
                ("bc_pk0", (C.c_double * MAXPH) * 6), ("bc_tank_p0", C.c_double * 6),
                ("bc_tank_T0", C.c_double * 6), ("bc_tank_frac", (C.c_double * MAXPH) * 6),
                ("bc_tank_massfrac", C.c_int * 6), ("ntr", C.c_int), ("bc_tr", (C.c_double * MAXTR) * 6),
                ("relax_pt", C.c_int), ("relax_ptmu", C.c_int), ("liq", C.c_int), ("vap", C.c_int),
                ("sref", C.c_double * MAXPH)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        dp = C.POINTER(C.c_double)
        L.orc_run.argtypes = [C.POINTER(OrcCfg), dp, dp, C.c_int, C.c_double, C.c_int, dp, dp, dp, C.c_int]
        L.orc_run.restype = C.c_int
        L.orc_run_from.argtypes = [C.POINTER(OrcCfg), dp, dp, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double,
                                   dp, dp, dp, C.c_int]
        L.orc_run_from.restype = C.c_int
        L.orc_limiter.argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_limiter.restype = C.c_double
        for name in ("orc_eos_energy", "orc_eos_sound"):
            getattr(L, name).argtypes = [C.POINTER(OrcEos), C.c_double, C.c_double, C.c_double]
            getattr(L, name).restype = C.c_double
        L.orc_eos_pressure.argtypes = [C.POINTER(OrcEos), C.c_double, C.c_double]
        L.orc_eos_pressure.restype = C.c_double
        L.orc_eos_isentrope.argtypes = [C.POINTER(OrcEos), C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_eos_isentrope.restype = C.c_double
        L.orc_eos_rho_pfinal.argtypes = [C.POINTER(OrcEos), C.c_double, C.c_double, C.c_double, dp, C.c_double]
        L.orc_eos_rho_pfinal.restype = C.c_double
        L.orc_relax_p.argtypes = [C.POINTER(OrcCfg), dp, dp, dp]
        L.orc_relax_p.restype = C.c_int
        L.orc_inflow_failure_step.argtypes = []
        L.orc_inflow_failure_step.restype = C.c_int
        L.orc_relax_failures.argtypes = [dp]
        L.orc_relax_failures.restype = C.c_long
        L.orc_hllc.argtypes = [C.POINTER(OrcCfg), dp, dp, C.c_double, C.c_double, dp, dp]
        L.orc_hllc.restype = None
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_eos(e) -> OrcEos:
    return OrcEos(EOS_TYPES[e["type"]], e["gamma"], e.get("pinf", 0.0), e["cv"], e.get("eref", 0.0))


def make_cfg(case) -> OrcCfg:
    cfg = OrcCfg()
    cfg.model = MODELS[case["model"]]
    cfg.nph = len(case["eos"])
    cfg.order = case["order"]
    for i, l in enumerate(case["limiters"]):
        cfg.lim[i] = LIMITERS[l]
    cfg.alpha_null = int(bool(case.get("alpha_null", False)))
    cfg.relax_p = int(bool(case.get("relax_p", False)))
    for d in range(3):
        cfg.n[d] = case["cells"][d]
        cfg.L[d] = case["lengths"][d]
    cfg.cfl = case["cfl"]
    for k, e in enumerate(case["eos"]):
        cfg.eos[k] = make_eos(e)
    for b, name in enumerate(case.get("bc", ["abs"] * 6)):
        cfg.bc[b] = BCS[name]
    for b, p0 in enumerate(case.get("bc_p0", [0.0] * 6)):
        cfg.bc_p0[b] = p0
    for b, inj in enumerate(case.get("bc_inj", [None] * 6)):
        if inj:
            cfg.bc_m0[b] = inj["m0"]
            for k in range(cfg.nph):
                cfg.bc_ak0[b][k] = inj["alpha"][k]
                cfg.bc_rhok0[b][k] = inj["rho"][k]
                cfg.bc_pk0[b][k] = inj["p"][k]
    for b, tk in enumerate(case.get("bc_tank", [None] * 6)):
        if tk:
            cfg.bc_tank_p0[b], cfg.bc_tank_T0[b] = tk["p0"], tk["T0"]
            frac = tk.get("massFrac") or tk.get("alpha") or [1.0]
            cfg.bc_tank_massfrac[b] = int("massFrac" in tk)
            for k, v in enumerate(frac):
                cfg.bc_tank_frac[b][k] = v
    cfg.ntr = len(case.get("transports", []))
    cfg.relax_pt = int(bool(case.get("relax_pt", False)))
    cfg.relax_ptmu = int(bool(case.get("relax_ptmu", False)))
    # liquid and vapour phases (EulerHomogeneous <flowModel liquid vapor>;
    # RelaxationPTMu always phase 0 and 1) and each EOS's entropyRef
    cfg.liq, cfg.vap = case.get("liq", 0), case.get("vap", 1)
    for k, e in enumerate(case["eos"]):
        cfg.sref[k] = e.get("sref", 0.0)
    for b, vals in enumerate(case.get("bc_tr", [None] * 6)):
        for t, v in enumerate(vals or []):
            cfg.bc_tr[b][t] = v
    return cfg


def ntransports(case) -> int:
    return len(case.get("transports", []))


def nfields(case) -> int:
    """per phase alpha, rho, p; u, v, w; E; the transports; ThermalEq, EulerHomogeneous: T last"""
    return 3 * len(case["eos"]) + 4 + ntransports(case) + (case["model"] in WITH_T)


def run(case, state, nsteps=None, t_final=None, max_steps=0, nthreads=0, log_dt=False, resume=None):
    """March ``state`` (field-major, shape (nfields, ncell) or (nfields-1, ncell)).
    ``resume`` = (t0, dt0) starts a resumed run (orc_run_from).

    Returns (state_out, steps, t, dt_next, dt_log or None).
    """
    cfg = make_cfg(case)
    ncell = int(np.prod(case["cells"]))
    nf = nfields(case)
    sin = np.zeros((nf, ncell))
    st = np.asarray(state, dtype=np.float64)
    sin[: st.shape[0]] = st
    sin = np.ascontiguousarray(sin)
    sout = np.zeros_like(sin)
    if nsteps is None and t_final is None:
        nsteps = case["steps"] if case["time_mode"] == "iterations" else None
        t_final = case.get("t_final", 0.0)
    t = np.zeros(1)
    dtn = np.zeros(1)
    cap = nsteps if nsteps else (max_steps if max_steps else 10_000_000)
    dtl = np.zeros(cap + 1) if log_dt else None
    t0, dt0 = resume if resume else (0.0, 1e-15)
    it = lib().orc_run_from(C.byref(cfg), _dp(sin), _dp(sout), int(nsteps or 0), float(t_final or 0.0),
                            int(max_steps), float(t0), float(dt0), _dp(t), _dp(dtn),
                            _dp(dtl) if log_dt else None, int(nthreads))
    if it < 0:
        raise RuntimeError(f"oracle error {it}")
    return sout, it, float(t[0]), float(dtn[0]), (dtl[:it] if log_dt else None)


def state_from_golden(case, d, prefix="init"):
    """Field-major state (alpha, rho, p per phase; u, v, w; with transports
    also E, left 0, and the transported scalars) from a golden .npz."""
    nph = len(case["eos"])
    ncell = int(np.prod(case["cells"]))
    ntr = ntransports(case)
    te = case["model"] in WITH_T
    out = np.zeros((3 * nph + 4 + ntr + te if (ntr or te) else 3 * nph + 3, ncell))
    if ntr:
        out[3 * nph + 4:3 * nph + 4 + ntr] = d[f"{prefix}_tr"]
    if te:
        out[3 * nph + 4] = d[f"{prefix}_T"]
    if case["model"] == "euler":
        out[0] = 1.0
        out[1] = d[f"{prefix}_rho"][0]
        out[2] = d[f"{prefix}_p"][0]
    else:
        for k in range(nph):
            out[3 * k] = d[f"{prefix}_alpha"][k]
            out[3 * k + 1] = d[f"{prefix}_rho"][k]
            out[3 * k + 2] = d[f"{prefix}_p"][k]
    out[3 * nph:3 * nph + 3] = d[f"{prefix}_u"]
    return out


def relax_failures():
    """(count, first inputs) of the relaxations of the last run that reached
    the 100-iteration limit; first inputs: per phase (alpha, rho, p) after
    verifyAndCorrectPhase, shape (MAXPH, 3)."""
    buf = np.zeros(3 * MAXPH)
    n = lib().orc_relax_failures(_dp(buf))
    return int(n), buf.reshape(MAXPH, 3)


def inflow_failure_step():
    """1-based step at which the last run stopped because an injection or tank
    Newton solve did not converge (the reference's process ends there); 0 if
    none did."""
    return int(lib().orc_inflow_failure_step())
