"""The allocation-free entry point (eco_run_noalloc) against the allocating
one (eco_create + eco_set_state + eco_run + eco_get_state).

eco_run_noalloc keeps every intermediate buffer in the library's static
device scratch; its result must equal the allocating path's bit for bit (the
same kernels, the same launches).  Several calls in a row share that scratch,
each smaller than the one before, so a call that read what an earlier call
left there would show.  The caller's device arrays come from hipMalloc in the
test itself (the caller allocates its inputs and outputs).  A solve larger
than the scratch is refused before any GPU work (no GPU needed for that
check).
"""
import ctypes as C

import numpy as np
import pytest

from conftest import initial_state, load_golden

from ecogen_amd import Context, workloads as W
from ecogen_amd import capi


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime, for the caller's arrays."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        return p.value

    def h2d(self, dst, arr):
        assert self.h.hipMemcpy(C.c_void_p(dst), arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1) == 0

    def d2h(self, arr, src):
        assert self.h.hipMemcpy(arr.ctypes.data_as(C.c_void_p), C.c_void_p(src), arr.nbytes, 2) == 0

    def free(self, p):
        assert self.h.hipFree(C.c_void_p(p)) == 0


def _host_layout(case, st0):
    nf = capi.nfields(case)
    ncell = int(np.prod(case["cells"]))
    s = np.zeros((nf, ncell))
    s[: st0.shape[0]] = st0
    return np.ascontiguousarray(s)


def _alloc_path(case, st0, nsteps, t_final):
    with Context(case) as ctx:
        ctx.set_state(st0)
        status = ctx.run(nsteps=nsteps) if nsteps > 0 else ctx.run(t_final=t_final)
        return ctx.get_state(), status


def _noalloc_path(hip, case, st0, nsteps, t_final):
    s_in = _host_layout(case, st0)
    s_in[-1] = np.nan  # E is ignored on input
    d_in, d_out = hip.malloc(s_in.nbytes), hip.malloc(s_in.nbytes)
    try:
        hip.h2d(d_in, s_in)
        hip.h2d(d_out, np.full_like(s_in, np.nan))  # every output value must be written
        status = capi.run_noalloc(case, d_in, d_out, nsteps=nsteps, t_final=t_final)
        out = np.empty_like(s_in)
        hip.d2h(out, d_out)
    finally:
        hip.free(d_in)
        hip.free(d_out)
    return out, status


def _cases():
    """Largest first: each call's scratch footprint is below the previous one's."""
    big, big0 = W.shock_cylinders_2d(n=160, ncyl=6, rmin=3.0, rmax=8.0, steps=40)
    c3, d3 = load_golden("c3_multip_3d_o2")          # MultiP + relaxation, 3-D, 16^3
    c1, d1 = load_golden("c1_kapila_fo")             # first order, time control
    mp, dm = load_golden("multip_1d_o2_relax")       # MultiP + relaxation, 1-D
    return [("cylinders 160^2 Kapila MUSCL, 40 steps", big, big0, 40, 0.0),
            ("c3_multip_3d_o2", c3, initial_state(c3, d3), c3["steps"], 0.0),
            ("c1_kapila_fo to 240 us", c1, initial_state(c1, d1), 0, c1["t_final"]),
            ("multip_1d_o2_relax", mp, initial_state(mp, dm), mp["steps"], 0.0)]


@pytest.mark.gpu
def test_noalloc_equals_allocating_path_on_calls_in_a_row():
    hip = _Hip()
    cases = _cases()
    sizes = [capi.noalloc_bytes(c) for _, c, _, _, _ in cases]
    assert all(0 < b <= capi.NOALLOC_SCRATCH_BYTES for b in sizes)
    assert sizes == sorted(sizes, reverse=True), sizes
    for name, case, st0, nsteps, t_final in cases:
        print(f"case {name}: eco_run_noalloc vs eco_run", flush=True)
        want, ws = _alloc_path(case, st0, nsteps, t_final)
        got, gs = _noalloc_path(hip, case, st0, nsteps, t_final)
        assert (gs.steps, gs.t, gs.dt_last, gs.dt_next, gs.flags, gs.done) == \
               (ws.steps, ws.t, ws.dt_last, ws.dt_next, ws.flags, ws.done), name
        for f in range(want.shape[0]):
            assert np.array_equal(got[f], want[f]), f"{name}: field {f} differs"


def test_noalloc_refuses_a_solve_beyond_the_scratch():
    """4096^2 in FP64 needs about 3.5 GB of scratch: refused with
    ECO_ERR_SCRATCH on the host, before any GPU call (dummy pointers)."""
    if not capi.LIB_PATH.exists():
        pytest.skip("libecogen_amd.so not built")
    case, _ = W.shock_cylinders_2d(n=4096, ncyl=1, steps=1)
    assert capi.noalloc_bytes(case) > capi.NOALLOC_SCRATCH_BYTES
    cfg = capi.make_config(case)
    st = capi.EcoStatus()
    rc = capi.lib().eco_run_noalloc(C.byref(cfg), C.c_void_p(16), C.c_void_p(4096), 1, 0.0, None, C.byref(st))
    assert rc == capi.ERR_SCRATCH
    assert "scratch" in capi.strerror(rc)
    # FP32 storage is not offered by the allocation-free form
    small, _ = W.shock_cylinders_2d(n=64, ncyl=1, rmin=2.0, rmax=4.0, steps=1)
    cfg = capi.make_config(dict(small, precision="fp32s"))
    assert capi.lib().eco_run_noalloc(C.byref(cfg), C.c_void_p(16), C.c_void_p(4096), 1, 0.0, None,
                                      C.byref(st)) == -2
