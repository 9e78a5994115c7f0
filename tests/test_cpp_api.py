"""C++ class API (include/ecogen.hpp): the descriptors lower to the C ABI
configuration with the reference's limiter cascade (host only), and a Run
driven through the class API gives the same solve as the C ABI (GPU)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

EXE = Path(__file__).resolve().parent / "cpp" / "api_check"


def _exe():
    if not EXE.exists():
        pytest.skip("tests/cpp/api_check not built (python -m ecogen_amd.build)")
    return EXE


def test_lowering_and_cascade():
    r = subprocess.run([str(_exe()), "lower"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK" in r.stdout


@pytest.mark.gpu
def test_run_matches_c_abi(tmp_path):
    out = tmp_path / "st.bin"
    r = subprocess.run([str(_exe()), "run", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(out, dtype=np.float64)
    steps, t, state = int(raw[0]), raw[1], raw[2:].reshape(10, 1000)
    from ecogen_amd import Context
    from ecogen_amd import workloads as W
    case, st = W.shock_tube_1d()
    with Context(case) as ctx:
        ctx.set_state(st)
        s = ctx.run(t_final=case["t_final"])
        ref = ctx.get_state()
    assert steps == s.steps and t == s.t
    assert np.array_equal(state, ref)


@pytest.mark.gpu
def test_resume_clock_matches_c_abi(tmp_path):
    """Run::resumeClock (eco_set_clock): 20 steps of the shock tube restarted
    at t = 1e-5 with a first dt of 3e-8, bit for bit the C ABI's, and the
    first step taken with that dt."""
    import os
    out = tmp_path / "st.bin"
    r = subprocess.run([str(_exe()), "run", str(out)], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, API_CHECK_RESUME="1"))
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(out, dtype=np.float64)
    steps, t, state = int(raw[0]), raw[1], raw[2:].reshape(10, 1000)
    from ecogen_amd import Context
    from ecogen_amd import workloads as W
    case, st = W.shock_tube_1d()
    with Context(case) as ctx:
        ctx.set_state(st)
        ctx.set_clock(1e-5, 3e-8)
        s1 = ctx.run(nsteps=1)
        assert s1.dt_last == 3e-8 and s1.t == 1e-5 + 3e-8
        s = ctx.run(nsteps=19)
        ref = ctx.get_state()
    assert steps == s.steps == 20 and t == s.t
    assert np.array_equal(state, ref)
