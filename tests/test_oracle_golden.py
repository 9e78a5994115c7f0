"""The CPU oracle against the reference's own outputs (golden vectors).

Each golden case was produced by the reference ECOGEN executable built from
source (tests/golden/make_golden.py); the oracle starts from the same t=0
state and must reproduce the final state, step count and final time.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from conftest import initial_state, compared_fields, golden_names, load_golden, rel_l1

import oracle as O


@pytest.mark.parametrize("name", golden_names())
def test_oracle_matches_reference(name):
    case, d = load_golden(name)
    st = initial_state(case, d)
    out, it, t, _, _ = O.run(case, st)
    info = d["info"]
    assert it == int(info[-1, 1])
    assert t == pytest.approx(info[-1, 2], rel=1e-15, abs=0)
    fin = O.state_from_golden(case, d, "final")
    for f, nm in compared_fields(case):
        if case["model"] == "kapila" and nm.startswith("p"):
            continue  # Kapila prints only the mixture pressure; checked below
        assert rel_l1(out[f], fin[f]) <= 1e-13, nm
    if case["model"] == "kapila":
        nph = len(case["eos"])
        pmix = sum(out[3 * k] * out[3 * k + 2] for k in range(nph))
        assert rel_l1(pmix, d["final_p_mix"]) <= 1e-13


STOP = sorted(p.stem for p in (Path(__file__).resolve().parent / "golden" / "stop").glob("*.npz"))


def load_stop(name):
    d = np.load(Path(__file__).resolve().parent / "golden" / "stop" / f"{name}.npz")
    return json.loads(bytes(d["case_json"]).decode()), d


@pytest.mark.parametrize("name", STOP)
def test_oracle_stops_where_the_reference_stops(name):
    """An injection whose Newton solve does not converge in 50 iterations ends
    the reference's process in that step (Errors::errorMessage after
    ModKapila.cpp:283 or ModEuler's solveRiemannInflow): its infoCalcul.out,
    written every iteration, ends at the last completed step.  The oracle
    stops in the next step (returning -3) and reproduces the times of the
    completed steps exactly."""
    case, d = load_stop(name)
    assert "not converged" in bytes(d["message"]).decode()
    info = d["info"]
    done = int(info[-1, 1])
    from conftest import initial_state_from_deck
    st = initial_state_from_deck(case)
    with pytest.raises(RuntimeError, match="-3"):
        O.run(case, st)
    assert O.inflow_failure_step() == done + 1
    _, it, t, _, _ = O.run(case, st, nsteps=done)
    assert it == done and t == info[-1, 2]
