"""Time-step control of eco_run: iteration mode, physical-time mode with the
float comparison of the reference (src/Run.cpp:239-244), the optional step
cap of the time mode, continuation after a finished run, and the
asynchronous launch form (steps past the target are no-ops on the device)."""
import numpy as np
import pytest

from ecogen_amd import Context
from ecogen_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tube():
    return W.shock_tube_1d()


def test_time_mode_cap(tube):
    case, st = tube
    with Context(case) as ctx:
        ctx.set_state(st)
        full = ctx.run(t_final=case["t_final"])
    assert full.done and full.t >= np.float32(case["t_final"])
    with Context(case) as ctx:
        ctx.set_state(st)
        s = ctx.run(t_final=case["t_final"], max_steps=100)
        assert s.steps == 100 and s.t < case["t_final"] and s.done
        a = ctx.get_state()
    with Context(case) as ctx:
        ctx.set_state(st)
        s2 = ctx.run(nsteps=100)
        b = ctx.get_state()
    assert s2.t == s.t and np.array_equal(a, b)
    with Context(case) as ctx:  # a cap above the step count changes nothing
        ctx.set_state(st)
        s3 = ctx.run(t_final=case["t_final"], max_steps=full.steps + 50)
    assert s3.steps == full.steps and s3.t == full.t


def test_continuation_and_async_noops(tube):
    case, st = tube
    with Context(case) as ctx:
        ctx.set_state(st)
        ctx.run(nsteps=40)
        s = ctx.run(nsteps=60)
        a = ctx.get_state()
    assert s.steps == 100
    with Context(case) as ctx:
        ctx.set_state(st)
        ctx.run(nsteps=100)
        b = ctx.get_state()
    assert np.array_equal(a, b)
    # asynchronous launches past a reached time target are no-ops
    with Context(case) as ctx:
        ctx.set_state(st)
        full = ctx.run(t_final=case["t_final"])
        ref = ctx.get_state()
    with Context(case) as ctx:
        ctx.set_state(st)
        ctx.run(t_final=case["t_final"])
        ctx.launch(25)
        s = ctx.sync()
        assert s.steps == full.steps and s.t == full.t
        assert np.array_equal(ctx.get_state(), ref)


def test_zero_copy_state_view(tube):
    """eco_state_device exposes the FP64 device state (set_state_device of
    that view reproduces it) and refuses the packed FP32-storage format."""
    case, st = tube
    with Context(case) as ctx:
        ctx.set_state(st)
        ctx.run(nsteps=10)
        ref = ctx.get_state()
        ptr, nf = ctx.state_device()
        assert ptr and nf == ref.shape[0] - 2  # device layout: 1-D, one velocity component
        with Context(case) as other:
            other.set_state_device(ptr)
            got = other.get_state()
    assert np.array_equal(got[:7], ref[:7])
    from ecogen_amd import workloads as W2
    c32, s32, _ = W2.riemann_case(1, 200, 10.0, 0.5, precision="fp32s")
    with Context(c32) as ctx:
        ctx.set_state(s32)
        ptr, nf = ctx.state_device()
        assert not ptr and nf == 0


@pytest.mark.parametrize("name", ["kapila_1d_fo_injection_stop", "euler_1d_fo_injection_stop"])
def test_run_stops_where_the_reference_stops(name):
    """An injection whose Newton solve does not converge in 50 iterations ends
    the reference's run in that step (tests/golden/stop: its last completed
    iteration and the message it printed); the oracle stops in the same step.
    The device flags it, makes every later step a no-op and eco_run returns
    ECO_ERR_NOT_CONVERGED, its status counting the (first-order) failing step
    as taken; the completed steps' times equal the reference's."""
    import json
    from pathlib import Path
    from conftest import initial_state_from_deck
    from ecogen_amd import EcoError
    from ecogen_amd.capi import ERR_NOT_CONVERGED
    import oracle as O
    d = np.load(Path(__file__).resolve().parent / "golden" / "stop" / f"{name}.npz")
    case = json.loads(bytes(d["case_json"]).decode())
    done = int(d["info"][-1, 1])
    st0 = initial_state_from_deck(case)
    print(f"case {name}: injection failure expected in step {done + 1}", flush=True)
    with Context(case) as ctx:
        ctx.set_state(st0)
        with pytest.raises(EcoError) as ei:
            ctx.run(nsteps=case["steps"])
        assert ei.value.code == ERR_NOT_CONVERGED
        s = ei.value.status
        assert s.flags & 4 and s.done and s.steps == done + 1
    with pytest.raises(RuntimeError):
        O.run(case, st0)
    assert O.inflow_failure_step() == done + 1
    with Context(case) as ctx:
        ctx.set_state(st0)
        s = ctx.run(nsteps=done)
        assert s.steps == done and s.t == d["info"][-1, 2] and s.flags == 0
