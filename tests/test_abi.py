"""The C-ABI library loads and exports every entry point the public headers
declare (no GPU needed: nothing here calls into device code), and the host
workload generators are deterministic."""
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADERS = [ROOT / "include" / "ecogen.h", ROOT / "include" / "ecogen_pool.h"]


def declared(header):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)  # drop comments
    return sorted(set(re.findall(r"\b(eco_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from ecogen_amd import capi
    if not capi.LIB_PATH.exists():
        pytest.skip("libecogen_amd.so not built (python -m ecogen_amd.build)")
    return capi.lib()


@pytest.mark.parametrize("header", HEADERS, ids=lambda h: h.name)
def test_exports_every_declared_symbol(lib, header):
    names = declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_lists_match_headers():
    from ecogen_amd.capi import EXPORTS
    from ecogen_amd.pool import POOL_EXPORTS
    assert sorted(EXPORTS) == declared(HEADERS[0])
    assert sorted(POOL_EXPORTS) == declared(HEADERS[1])


def test_version_and_strerror(lib):
    assert b"gfx950" in lib.eco_version()
    assert lib.eco_strerror(0) == b"ok"
    assert lib.eco_strerror(-2).startswith(b"configuration not supported")


def test_case_bytes_without_device(lib):
    """eco_pool_case_bytes is pure host arithmetic: FP32 storage halves the
    state bytes of the FP64 form (two time levels + U*, plus distance tables)."""
    from ecogen_amd import workloads as W
    from ecogen_amd.pool import Pool
    c32, _, _ = W.riemann_case(2, 100, 10.0, 0.5, precision="fp32s")
    c64 = dict(c32, precision="fp64")
    b32, b64 = Pool.case_bytes(c32), Pool.case_bytes(c64)
    nc = 100 * 100
    assert b32 >= nc * (2 * 9 + 8) * 4
    assert b64 >= nc * (2 * 9 + 8) * 8
    assert b64 > 1.9 * b32


def test_batch_specs_deterministic_and_distributed():
    from ecogen_amd import workloads as W
    a = W.batch_specs(2048)
    assert a == W.batch_specs(2048)
    d1 = [s for s in a if s[0] == 1]
    d2 = [s for s in a if s[0] == 2]
    assert 0.65 < len(d1) / len(a) < 0.75
    assert all(200 <= s[1] <= 50000 for s in d1) and all(64 <= s[1] <= 768 for s in d2)
    assert all(10.0 <= s[2] <= 1e4 and 0.2 <= s[3] <= 0.8 for s in a)
    # log-uniform sizes: median of the 1-D sizes near the geometric mean sqrt(200*50000) = 3162
    assert 2000 < np.median([s[1] for s in d1]) < 5000


def test_request_stream_rate_scaling():
    from ecogen_amd import workloads as W
    a = W.request_stream(50.0, 200)
    b = W.request_stream(800.0, 200)
    assert [x[1:] for x in a] == [x[1:] for x in b]
    ta = np.array([x[0] for x in a])
    tb = np.array([x[0] for x in b])
    assert np.allclose(ta * 50.0, tb * 800.0)
    assert np.all(np.diff(ta) > 0)
    # mean rate within 25 % of lambda over 200 arrivals
    assert 0.75 < (200 / ta[-1]) / 50.0 < 1.25
    kinds = [x[1] for x in a]
    assert 0 < kinds.count("bubbles") < 30


def test_riemann_case_t_final():
    """t_final = 0.3 L / a_max with a_max the frozen sound speed of the
    high-pressure water state (stiffened gas, fluid at rest)."""
    from ecogen_amd import workloads as W
    case, st, tf = W.riemann_case(1, 1000, 100.0, 0.5)
    c_w = np.sqrt(4.4 * (1e7 + 6e8) / 1000.0)
    assert abs(tf - 0.3 / c_w) / tf < 1e-3  # water mass fraction ~1 - 5e-8
    assert st.shape == (9, 1000)
    assert st[2][0] == 1e7 and st[2][-1] == 1e5


def test_boundary_configs_are_validated_without_device(lib):
    """eco_create checks a configuration before any HIP call: an injection
    side needs a nonzero flow, positive densities and fractions summing to 1
    (BoundCondInj.cpp:46-150), a tank positive p0 and T0 and fractions summing
    to 1 (BoundCondTank.cpp:46-144); both run in FP64 only."""
    import ctypes as C
    from ecogen_amd import capi
    from cases import CASES
    case = CASES["kapila_1d_o2_injection"]

    def create(c):
        ctx = C.c_void_p()
        return lib.eco_create(C.byref(capi.make_config(c)), None, C.byref(ctx))
    inj = case["bc_inj"][1]
    assert create(dict(case, precision="fp32s")) == -2
    bad_sum = [None, dict(inj, alpha=[0.5, 0.4])] + [None] * 4
    assert create(dict(case, bc_inj=bad_sum)) == -1
    no_flow = [None, dict(inj, m0=0.0)] + [None] * 4
    assert create(dict(case, bc_inj=no_flow)) == -1
    no_rho = [None, dict(inj, rho=[50.0, 0.0])] + [None] * 4
    assert create(dict(case, bc_inj=no_rho)) == -1
    tcase = CASES["multip_1d_o2_relax_tank"]
    tk = tcase["bc_tank"][1]
    assert create(dict(tcase, precision="fp32s")) == -2
    assert create(dict(tcase, bc_tank=[None, dict(tk, massFrac=[0.5, 0.6])] + [None] * 4)) == -1
    assert create(dict(tcase, bc_tank=[None, dict(tk, T0=0.0)] + [None] * 4)) == -1


def test_phase_counts_are_validated_without_device(lib):
    """Three and four phases run in FP64 single solves; FP32 storage and
    THINC (a two-phase reconstruction) refuse them; five phases are invalid."""
    import ctypes as C
    from ecogen_amd import capi
    from cases import CASES

    def create(c):
        ctx = C.c_void_p()
        return lib.eco_create(C.byref(capi.make_config(c)), None, C.byref(ctx))
    c3 = CASES["kapila_1d_o2_3ph"]
    assert create(dict(c3, precision="fp32s")) == -2
    assert create(dict(c3, limiters=["vanleer", "vanleer", "thinc", "thinc"])) == -2
    c4 = CASES["kapila_1d_o2_4ph"]
    assert create(dict(c4, precision="fp32s")) == -2
    cfg = capi.make_config(c4)
    cfg.nphase = 5  # eco_config holds four EOS
    ctx = C.c_void_p()
    assert lib.eco_create(C.byref(cfg), None, C.byref(ctx)) == -1


def test_transport_configs_are_validated_without_device(lib):
    """Transported scalars: up to four, Kapila and MultiP in FP64 single
    solves.  The reference's Euler model has no transport Riemann solver
    (Model::solveRiemannTransportIntern ends the run, Model.h:127), so an
    Euler configuration with transports is invalid; FP32 storage, pools and
    the allocation-free entry point refuse them."""
    import ctypes as C
    from ecogen_amd import capi
    from cases import CASES

    def create(c):
        ctx = C.c_void_p()
        return lib.eco_create(C.byref(capi.make_config(c)), None, C.byref(ctx))
    c = CASES["kapila_1d_o2_transport_thinc"]
    assert create(dict(c, precision="fp32s")) == -2
    assert create(dict(CASES["euler_1d_o2"], transports=["color"])) == -1
    assert create(dict(c, transports=["a", "b", "c", "d", "e"])) == -1
    assert lib.eco_noalloc_bytes(C.byref(capi.make_config(dict(CASES["euler_1d_o2"], transports=["c"])))) == 0
    buf = (C.c_double * 4)()
    st = capi.EcoStatus()
    rc = lib.eco_run_noalloc(C.byref(capi.make_config(c)), C.addressof(buf), C.addressof(buf) + 16,
                             5, C.c_double(0.0), None, C.byref(st))
    assert rc == -2
    assert capi.nfields(c) == 3 * 2 + 4 + 2


def test_thermaleq_configs_are_validated_without_device(lib):
    """The thermal-equilibrium model: two phases (MixThermalEq::computePressure
    refuses more), no transport or injection solver (the reference ends the
    run there), the tank and FP32 storage not provided; its host state carries
    the temperature after E."""
    import ctypes as C
    from ecogen_amd import capi
    from cases import CASES

    def create(c):
        ctx = C.c_void_p()
        return lib.eco_create(C.byref(capi.make_config(c)), None, C.byref(ctx))
    te = CASES["thermaleq_1d_o2_vanleer"]
    assert create(dict(te, precision="fp32s")) == -2
    assert create(dict(te, transports=["c"])) == -1
    assert create(dict(te, eos=te["eos"] + [te["eos"][0]])) == -1
    inj = CASES["kapila_1d_o2_injection"]
    assert create(dict(te, bc=inj["bc"], bc_inj=inj["bc_inj"])) == -1
    tk = CASES["kapila_1d_o2_tank"]
    assert create(dict(te, bc=tk["bc"], bc_tank=tk["bc_tank"])) == -2
    assert capi.nfields(te) == 3 * 2 + 5
    assert lib.eco_noalloc_bytes(C.byref(capi.make_config(te))) == 0
