"""ctypes binding of the C ABI (include/ecogen.h) of ``libecogen_amd.so``.

The library is built in-tree by ``ecogen_amd.build``; this module only loads
it.  There is no fallback: if the shared library is missing or a call fails,
an exception is raised.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libecogen_amd.so"

MODEL = {"euler": 0, "kapila": 1, "multip": 2, "thermaleq": 3, "eulerhomogeneous": 4}
# models whose host state ends with the mixture temperature
WITH_T = ("thermaleq", "eulerhomogeneous")
LIMITER = {"none": 0, "minmod": 1, "vanleer": 2, "vanalbada": 3, "superbee": 4, "mc": 5, "thinc": 6}
EOS_TYPE = {"IG": 0, "SG": 1}
PRECISION = {"fp64": 0, "fp32s": 1}
# per side X-, X+, Y-, Y+, Z-, Z+; all but abs: FP64 only.  Outflow pressures in
# case["bc_p0"]; injections in case["bc_inj"][side] = {"m0": mass flow into the
# domain, "alpha", "rho", "p": the injected fluid per phase}; tanks in
# case["bc_tank"][side] = {"p0", "T0", and per phase "alpha" or "massFrac"}
BC = {"abs": 0, "wall": 1, "outflow": 2, "injection": 3, "tank": 4}
# transported scalars: case["transports"] = their names, each state's "tr" =
# values in that order, case["bc_tr"][side] = the values entering through an
# outflow, injection or tank side; stored after E (ECO_NFIELDS_TR)
MAX_TRANSPORTS = 4

WARN_RELAX_NOT_CONVERGED = 1
WARN_NAN = 2
WARN_INFLOW_NOT_CONVERGED = 4

EXPORTS = ["eco_create", "eco_set_state", "eco_set_state_device", "eco_set_clock", "eco_get_state", "eco_step", "eco_run",
           "eco_launch_steps", "eco_sync", "eco_time_steps", "eco_copy_calibration", "eco_state_device", "eco_destroy", "eco_strerror", "eco_version",
           "eco_run_noalloc", "eco_noalloc_bytes"]

ERR_SCRATCH = -8            # eco_run_noalloc: the solve exceeds the static scratch
ERR_NOT_CONVERGED = -9      # an injection or tank solve stopped the run (the reference exits there)
NOALLOC_SCRATCH_BYTES = 256 << 20  # ECO_NOALLOC_SCRATCH_BYTES


class EcoEos(C.Structure):
    _fields_ = [("type", C.c_int), ("gamma", C.c_double), ("pinf", C.c_double), ("cv", C.c_double),
                ("eref", C.c_double)]


class EcoConfig(C.Structure):
    _fields_ = [("model", C.c_int), ("nphase", C.c_int), ("order", C.c_int), ("lim", C.c_int * 4),
                ("alpha_null", C.c_int), ("relax_p", C.c_int), ("precision", C.c_int), ("n", C.c_int * 3),
                ("L", C.c_double * 3), ("cfl", C.c_double), ("bc", C.c_int * 6), ("eos", EcoEos * 4),
                ("bc_p0", C.c_double * 6), ("bc_inj_m0", C.c_double * 6),
                ("bc_inj_alpha", (C.c_double * 4) * 6), ("bc_inj_rho", (C.c_double * 4) * 6),
                ("bc_inj_p", (C.c_double * 4) * 6), ("bc_tank_p0", C.c_double * 6),
                ("bc_tank_T0", C.c_double * 6), ("bc_tank_frac", (C.c_double * 4) * 6),
                ("bc_tank_massfrac", C.c_int * 6), ("ntransport", C.c_int),
                ("bc_tr", (C.c_double * MAX_TRANSPORTS) * 6), ("relax_pt", C.c_int),
                ("relax_ptmu", C.c_int), ("liq", C.c_int), ("vap", C.c_int), ("sref", C.c_double * 4)]


class EcoStatus(C.Structure):
    _fields_ = [("steps", C.c_int), ("t", C.c_double), ("dt_last", C.c_double), ("dt_next", C.c_double),
                ("flags", C.c_uint32), ("done", C.c_int)]


class EcoError(RuntimeError):
    def __init__(self, code: int, where: str):
        super().__init__(f"{where}: {strerror(code)} ({code})")
        self.code = code


_lib = None


def lib():
    """Load the device library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} not built; run `python -m ecogen_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.eco_create.argtypes = [C.POINTER(EcoConfig), vp, C.POINTER(vp)]
        L.eco_create.restype = C.c_int
        L.eco_set_state.argtypes = [vp, dp]
        L.eco_set_state.restype = C.c_int
        L.eco_set_state_device.argtypes = [vp, vp]
        L.eco_set_state_device.restype = C.c_int
        L.eco_set_clock.argtypes = [vp, C.c_double, C.c_double]
        L.eco_set_clock.restype = C.c_int
        L.eco_get_state.argtypes = [vp, dp]
        L.eco_get_state.restype = C.c_int
        L.eco_step.argtypes = [vp, C.POINTER(EcoStatus)]
        L.eco_step.restype = C.c_int
        L.eco_run.argtypes = [vp, C.c_int, C.c_double, C.POINTER(EcoStatus)]
        L.eco_run.restype = C.c_int
        L.eco_launch_steps.argtypes = [vp, C.c_int]
        L.eco_launch_steps.restype = C.c_int
        L.eco_sync.argtypes = [vp, C.POINTER(EcoStatus)]
        L.eco_sync.restype = C.c_int
        L.eco_time_steps.argtypes = [vp, C.c_int, dp, C.POINTER(C.c_int), C.POINTER(EcoStatus)]
        L.eco_time_steps.restype = C.c_int
        L.eco_copy_calibration.argtypes = [vp, vp, C.c_ulonglong, vp]
        L.eco_copy_calibration.restype = C.c_int
        L.eco_state_device.argtypes = [vp, C.POINTER(C.c_int)]
        L.eco_state_device.restype = vp
        L.eco_destroy.argtypes = [vp]
        L.eco_destroy.restype = None
        L.eco_strerror.argtypes = [C.c_int]
        L.eco_strerror.restype = C.c_char_p
        L.eco_version.argtypes = []
        L.eco_version.restype = C.c_char_p
        L.eco_run_noalloc.argtypes = [C.POINTER(EcoConfig), vp, vp, C.c_int, C.c_double, vp, C.POINTER(EcoStatus)]
        L.eco_run_noalloc.restype = C.c_int
        L.eco_noalloc_bytes.argtypes = [C.POINTER(EcoConfig)]
        L.eco_noalloc_bytes.restype = C.c_ulonglong
        _lib = L
    return _lib


def strerror(code: int) -> str:
    try:
        return lib().eco_strerror(code).decode()
    except Exception:  # library not loadable
        return f"error {code}"


def _check(rc: int, where: str):
    if rc != 0:
        raise EcoError(rc, where)


def make_config(case) -> EcoConfig:
    """Lower a case description (dict, see tests/golden/cases.py) to eco_config."""
    c = EcoConfig()
    c.model = MODEL[case["model"]]
    c.nphase = len(case["eos"])
    c.order = case["order"]
    for i, l in enumerate(case["limiters"]):
        c.lim[i] = LIMITER[l]
    c.alpha_null = int(bool(case.get("alpha_null", False)))
    c.relax_p = int(bool(case.get("relax_p", False)))
    c.precision = PRECISION[case.get("precision", "fp64")]
    for d in range(3):
        c.n[d] = case["cells"][d]
        c.L[d] = case["lengths"][d]
    c.cfl = case["cfl"]
    for b, name in enumerate(case.get("bc", ["abs"] * 6)):
        c.bc[b] = BC[name]
    for b, p0 in enumerate(case.get("bc_p0", [0.0] * 6)):
        c.bc_p0[b] = p0
    for b, inj in enumerate(case.get("bc_inj") or [None] * 6):
        if inj:
            c.bc_inj_m0[b] = inj["m0"]
            for k in range(len(inj["rho"])):
                c.bc_inj_alpha[b][k] = inj["alpha"][k]
                c.bc_inj_rho[b][k] = inj["rho"][k]
                c.bc_inj_p[b][k] = inj["p"][k]
    for b, tk in enumerate(case.get("bc_tank") or [None] * 6):
        if tk:
            c.bc_tank_p0[b], c.bc_tank_T0[b] = tk["p0"], tk["T0"]
            c.bc_tank_massfrac[b] = int("massFrac" in tk)
            for k, v in enumerate(tk.get("massFrac") or tk.get("alpha") or [1.0]):
                c.bc_tank_frac[b][k] = v
    for k, e in enumerate(case["eos"]):
        c.eos[k] = EcoEos(EOS_TYPE[e["type"]], e["gamma"], e.get("pinf", 0.0), e["cv"], e.get("eref", 0.0))
    c.ntransport = len(case.get("transports", []))
    c.relax_pt = int(bool(case.get("relax_pt", False)))
    c.relax_ptmu = int(bool(case.get("relax_ptmu", False)))
    c.liq, c.vap = case.get("liq", 0), case.get("vap", 1)
    for k, e in enumerate(case["eos"]):
        c.sref[k] = e.get("sref", 0.0)
    for b, vals in enumerate(case.get("bc_tr") or [None] * 6):
        for t, v in enumerate(vals or []):
            c.bc_tr[b][t] = v
    return c


def nfields(case) -> int:
    """Fields of a host state: per phase alpha, rho, p; u, v, w; E; the
    transports; for the thermal-equilibrium and homogeneous liquid-vapour
    models the mixture temperature last."""
    return 3 * len(case["eos"]) + 4 + len(case.get("transports", [])) + (case["model"] in WITH_T)


class Context:
    """One device solve behind the C ABI."""

    def __init__(self, case, stream=None):
        self.case = case
        self.cfg = make_config(case)
        self.ncell = int(np.prod(case["cells"]))
        self.nf = nfields(case)
        h = C.c_void_p()
        _check(lib().eco_create(C.byref(self.cfg), stream, C.byref(h)), "eco_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().eco_destroy(self._h)
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

    def set_state(self, state):
        s = np.zeros((self.nf, self.ncell))
        st = np.asarray(state, dtype=np.float64)
        s[: st.shape[0]] = st.reshape(st.shape[0], self.ncell)
        s = np.ascontiguousarray(s)
        _check(lib().eco_set_state(self._h, s.ctypes.data_as(C.POINTER(C.c_double))), "eco_set_state")

    def set_clock(self, t: float, dt: float):
        """Resume at physical time t with dt for the first step (eco_set_clock)."""
        _check(lib().eco_set_clock(self._h, C.c_double(t), C.c_double(dt)), "eco_set_clock")

    def set_state_device(self, ptr: int):
        _check(lib().eco_set_state_device(self._h, C.c_void_p(ptr)), "eco_set_state_device")

    def get_state(self):
        s = np.zeros((self.nf, self.ncell))
        _check(lib().eco_get_state(self._h, s.ctypes.data_as(C.POINTER(C.c_double))), "eco_get_state")
        return s

    def step(self) -> EcoStatus:
        st = EcoStatus()
        _check(lib().eco_step(self._h, C.byref(st)), "eco_step")
        return st

    def run(self, nsteps: int = 0, t_final: float = 0.0, max_steps: int = 0) -> EcoStatus:
        """nsteps > 0: exactly nsteps steps; otherwise until t >= t_final (as
        float), stopping after max_steps steps if max_steps > 0.  A run that
        an injection or tank solve stopped raises EcoError (code
        ERR_NOT_CONVERGED) carrying the status in ``.status``."""
        if nsteps <= 0 and max_steps > 0:
            nsteps = -int(max_steps)
        st = EcoStatus()
        rc = lib().eco_run(self._h, int(nsteps), float(t_final), C.byref(st))
        if rc == ERR_NOT_CONVERGED:
            err = EcoError(rc, "eco_run")
            err.status = st
            raise err
        _check(rc, "eco_run")
        return st

    def launch(self, nsteps: int):
        _check(lib().eco_launch_steps(self._h, int(nsteps)), "eco_launch_steps")

    def sync(self) -> EcoStatus:
        st = EcoStatus()
        _check(lib().eco_sync(self._h, C.byref(st)), "eco_sync")
        return st

    def time_steps(self, nsteps: int):
        """Run nsteps steps with HIP events around each stage kernel.

        Returns (status, stage_ms[3], launches[3]) for the first-order,
        predictor and corrector kernels."""
        ms = np.zeros(3)
        cnt = (C.c_int * 3)()
        st = EcoStatus()
        _check(lib().eco_time_steps(self._h, int(nsteps), ms.ctypes.data_as(C.POINTER(C.c_double)), cnt,
                                    C.byref(st)), "eco_time_steps")
        return st, ms, list(cnt)

    def state_device(self):
        nf = C.c_int()
        p = lib().eco_state_device(self._h, C.byref(nf))
        return p, nf.value


def noalloc_bytes(case) -> int:
    """Scratch bytes eco_run_noalloc needs for the case (0: invalid case)."""
    cfg = make_config(case)
    return int(lib().eco_noalloc_bytes(C.byref(cfg)))


def run_noalloc(case, dev_in: int, dev_out: int, nsteps: int = 0, t_final: float = 0.0, stream=None) -> EcoStatus:
    """eco_run_noalloc from device array ``dev_in`` to ``dev_out`` (both in the
    ecogen.h layout, 3*nphase+4 fields of ncell doubles): the allocation-free
    form of a solve (static device scratch of NOALLOC_SCRATCH_BYTES)."""
    cfg = make_config(case)
    st = EcoStatus()
    _check(lib().eco_run_noalloc(C.byref(cfg), C.c_void_p(dev_in), C.c_void_p(dev_out), int(nsteps), float(t_final),
                                 stream, C.byref(st)), "eco_run_noalloc")
    return st
