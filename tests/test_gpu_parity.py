"""HIP path against the oracle and the reference goldens (FP64).

Every golden case (all models, first order and MUSCL, every limiter, THINC,
alphaNull, relaxation, 1-D/2-D/3-D, every boundary type, three and four
phases) runs through the C ABI on the GPU from the reference's t=0 state.
The device computes every operation with the reference's rounding (IEEE
+, -, *, /, sqrt without contraction; pow and exp as the C library computes
them), so its final state must equal the oracle's bit for bit in every field
(per-phase alpha_k, rho_k, p_k, velocity, E), with the same number of steps
and the same final time.  The liquid-vapour saturation laws (EulerHomogeneous,
RelaxationPTMu) also take the C library's log, which misrounds about one
argument in 10^4 to 10^5 where the device's log is correctly rounded: there
every field is held to 1e-11 relative L1 of the oracle and of the reference
(a Newton solve stopping at |f| <= 1e-10 may stop an iterate apart) and the
final time to 1e-12.  Against the reference's own printed output the
bar is 1e-12 relative L1 in mixture density, velocity, mixture pressure and
per-phase volume fraction (the oracle reproduces those files exactly).
"""
import numpy as np
import pytest

from conftest import initial_state, golden_names, load_golden, parity_quantities, rel_l1

import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12        # FP64 parity bar against the reference's printed output, relative L1
TOL_SAT = 1e-11    # the same with the liquid-vapour saturation laws (the C library's log, see below)


def saturation_physics(case):
    """Cases whose step solves the liquid-vapour saturation laws
    (EulerHomogeneous, RelaxationPTMu)."""
    return case["model"] == "eulerhomogeneous" or bool(case.get("relax_ptmu"))


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
    st0 = initial_state(case, d)
    ref, it, t, _, _ = O.run(case, st0)
    out, status = _run_gpu(case, st0)
    assert status.steps == it
    assert status.flags & 2 == 0
    if saturation_physics(case):
        # the saturation temperature takes the C library's log, which
        # misrounds about one argument in 10^4 to 10^5 where the device's is
        # correctly rounded (tests/test_math.py); a Newton solve stopping at
        # |f| <= 1e-10 may then stop an iterate earlier or later
        assert status.t == pytest.approx(t, rel=1e-12, abs=0)
        for f in range(out.shape[0]):
            e = rel_l1(out[f], ref[f])
            assert e <= TOL_SAT, f"field {f}: rel L1 vs oracle {e:.3e}"
    else:
        assert status.t == t
        for f in range(out.shape[0]):
            nd = int(np.count_nonzero(out[f] != ref[f]))
            assert nd == 0, f"field {f}: {nd} cells differ from the oracle (max {np.abs(out[f] - ref[f]).max():.3e})"
    fin = O.state_from_golden(case, d, "final")
    got, want = parity_quantities(case, out), parity_quantities(case, ref)
    for nm in got:
        e = rel_l1(got[nm], want[nm])
        assert e <= (TOL_SAT if saturation_physics(case) else TOL), f"{nm}: rel L1 vs oracle {e:.3e}"
    ref_final = parity_quantities(case, fin)
    for nm in got:
        if case["model"] == "multip" and nm == "p" and case.get("relax_p"):
            continue  # the reference prints p* of the relaxation, not sum alpha_k p_k
        if case["model"] == "kapila" and nm == "p":
            e2 = rel_l1(got[nm], d["final_p_mix"])
        else:
            e2 = rel_l1(got[nm], ref_final[nm])
        assert e2 <= (TOL_SAT if saturation_physics(case) else TOL), f"{nm}: rel L1 vs reference {e2:.3e}"


def test_walls_are_refused_in_fp32_storage():
    """The FP32-storage kernels keep transmissive boundaries: a wall there is
    refused, not run as a transmissive boundary."""
    from ecogen_amd import Context, EcoError
    case, _ = load_golden("kapila_1d_o2_wall")
    with pytest.raises(EcoError):
        Context(dict(case, precision="fp32s"))
    Context(case).close()
