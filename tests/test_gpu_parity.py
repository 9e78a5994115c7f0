"""HIP path against the oracle and the reference goldens (FP64, <= 1e-12 relative L1).

Every golden case (all models, first order and MUSCL, every limiter, THINC,
alphaNull, relaxation, 1-D/2-D/3-D) runs through the C ABI on the GPU from
the reference's t=0 state; the final state must match the oracle and the
reference's output within 1e-12 relative L1 per field, with the same number
of steps and the same final time.
"""
import numpy as np
import pytest

from conftest import compared_fields, golden_names, load_golden, rel_l1

import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12  # FP64 parity bar, relative L1 per field


def _run_gpu(case, state):
    from ecogen_amd import Context
    with Context(case) as ctx:
        ctx.set_state(state)
        if case["time_mode"] == "iterations":
            st = ctx.run(nsteps=case["steps"])
        else:
            st = ctx.run(t_final=case["t_final"])
        return ctx.get_state(), st


@pytest.mark.parametrize("name", golden_names())
def test_gpu_matches_oracle_and_reference(name):
    case, d = load_golden(name)
    print(f"case {name}: {case['model']} order={case['order']} cells={case['cells']}", flush=True)
    st0 = O.state_from_golden(case, d)
    ref, it, t, _, _ = O.run(case, st0)
    out, status = _run_gpu(case, st0)
    assert status.steps == it
    assert status.t == pytest.approx(t, rel=1e-12)
    assert status.flags & 2 == 0
    fin = O.state_from_golden(case, d, "final")
    for f, nm in compared_fields(case):
        e = rel_l1(out[f], ref[f])
        assert e <= TOL, f"{nm}: rel L1 vs oracle {e:.3e}"
        if not (case["model"] == "kapila" and nm.startswith("p")):
            e2 = rel_l1(out[f], fin[f])
            assert e2 <= TOL, f"{nm}: rel L1 vs reference {e2:.3e}"
