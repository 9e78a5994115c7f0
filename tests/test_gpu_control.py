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
