"""Shared pytest configuration: the ``gpu`` marker and golden-case helpers."""
import json
import os
import sys
from pathlib import Path

# The oracle is built with OpenMP.  Under pytest-xdist every worker would
# start one thread per core and the spinning teams oversubscribe the host
# (a 0.1 s oracle run took minutes): one thread per worker there.
if os.environ.get("PYTEST_XDIST_WORKER"):
    os.environ["OMP_NUM_THREADS"] = "1"

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(GOLDEN))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def golden_names():
    return sorted(p.stem for p in GOLDEN.glob("*.npz"))


def load_golden(name):
    d = np.load(GOLDEN / f"{name}.npz")
    case = json.loads(bytes(d["case_json"]).decode())
    return case, d


def initial_state(case, d):
    """The t = 0 state of a golden: the reference's printed fields, except a
    Kapila deck's phase pressures, which come from the deck itself (the
    reference prints only sum alpha_k p_k, which does not give p_k back
    exactly when the alpha_k do not sum to 1 in floating point, as with three
    phases; the deck reader reproduces every printed field, tests/test_deck.py)."""
    import tempfile
    import oracle as O
    st = O.state_from_golden(case, d)
    if case["model"] == "kapila":
        from deck import write_deck
        from ecogen_amd import deck
        with tempfile.TemporaryDirectory() as tmp:
            write_deck(dict(case, name="t0"), Path(tmp) / "run")
            _, s0 = deck.load(Path(tmp) / "run")
        for k in range(len(case["eos"])):
            st[3 * k + 2] = s0[3 * k + 2]
    return st


def initial_state_from_deck(case):
    """The t = 0 state the deck reader builds from a case's deck (the
    reference's geometric domains on its cell centres)."""
    import tempfile
    from deck import write_deck
    from ecogen_amd import deck
    with tempfile.TemporaryDirectory() as tmp:
        write_deck(dict(case, name="t0"), Path(tmp) / "run")
        _, s0 = deck.load(Path(tmp) / "run")
    return s0


def rel_l1(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).sum()
    if den == 0:
        return float(np.abs(a - b).max())
    return float(np.abs(a - b).sum() / den)


def parity_quantities(case, state):
    """The quantities the parity bars are stated on: mixture density, velocity
    (one vector field: the L1 norm sums the components), mixture pressure and
    per-phase volume fraction (dict name -> array)."""
    nph = len(case["eos"])
    dim = sum(1 for c in case["cells"] if c > 1)
    out = {}
    if case["model"] == "euler":
        out["rho"] = state[1]
        out["p"] = state[2]
    elif case["model"] in ("thermaleq", "eulerhomogeneous"):
        # one pressure and temperature (the temperature after E)
        out["rho"] = sum(state[3 * k] * state[3 * k + 1] for k in range(nph))
        out["p"] = state[2]
        out["T"] = state[3 * nph + 4]
        for k in range(nph):
            out[f"alpha{k}"] = state[3 * k]
    else:
        out["rho"] = sum(state[3 * k] * state[3 * k + 1] for k in range(nph))
        out["p"] = sum(state[3 * k] * state[3 * k + 2] for k in range(nph))
        for k in range(nph):
            out[f"alpha{k}"] = state[3 * k]
    out["velocity"] = np.concatenate([state[3 * nph + d] for d in range(dim)])
    # transported scalars (stored after E)
    for t in range(len(case.get("transports", []))):
        out[f"tr{t}"] = state[3 * nph + 4 + t]
    return out


def phase_quantities(case, state):
    """Per-phase density and pressure (multiphase models)."""
    nph = len(case["eos"])
    if case["model"] == "euler":
        return {}
    out = {}
    for k in range(nph):
        out[f"rho{k}"] = state[3 * k + 1]
        out[f"p{k}"] = state[3 * k + 2]
    return out


def compared_fields(case):
    """(field index in the ecogen.h layout, name) pairs compared at parity."""
    nph = len(case["eos"])
    dim = sum(1 for c in case["cells"] if c > 1)
    out = []
    if case["model"] == "euler":
        out += [(1, "rho"), (2, "p")]
    else:
        for k in range(nph):
            out += [(3 * k, f"alpha{k}"), (3 * k + 1, f"rho{k}"), (3 * k + 2, f"p{k}")]
    for d in range(dim):
        out.append((3 * nph + d, "uvw"[d]))
    # transported scalars after E; the thermal-equilibrium temperature
    for t in range(len(case.get("transports", []))):
        out.append((3 * nph + 4 + t, f"tr{t}"))
    if case["model"] in ("thermaleq", "eulerhomogeneous"):
        out.append((3 * nph + 4, "T"))
    return out


def conserved_totals(case, state):
    """Per-phase mass, mixture momentum, sum rho|u| per axis and mixture total
    energy of a two-phase state, each summed over the cells times the cell volume."""
    nph = 2
    dV = float(np.prod([case["lengths"][d] / case["cells"][d] for d in range(3)]))
    rho = state[0] * state[1] + state[3] * state[4]
    mass = np.array([(state[3 * k] * state[3 * k + 1]).sum() for k in range(nph)]) * dV
    mom = np.array([(rho * state[3 * nph + d]).sum() for d in range(3)]) * dV
    amom = np.array([(rho * np.abs(state[3 * nph + d])).sum() for d in range(3)]) * dV
    energy = float((rho * state[3 * nph + 3]).sum()) * dV
    return mass, mom, amom, energy


def shock_inflow_errors(case, a, t1, b, t2):
    """Balance of the planar-shock inflow problem (post-shock state on the
    left boundary column, supersonic inflow, air at rest at 1e5 Pa on the
    right, nothing yet at the other boundaries) between states ``a`` at time
    ``t1`` and ``b`` at ``t2``: relative errors of per-phase mass, mixture
    energy and x-momentum against what enters through x = 0, and the y-momentum
    over sum rho|u|."""
    n = case["cells"][0]
    s = a[:, 0]
    # the boundary column keeps the post-shock state (alpha, rho, p, u; rows
    # away from the corners): its physical flux is what crosses x = 0
    for st in (a, b):
        assert np.allclose(st[:7, 2 * n:-2 * n:n], s[:7, None], rtol=1e-13, atol=0.0)
    rho_s = s[0] * s[1] + s[3] * s[4]
    u_s, p_s, E_s = s[6], s[0] * s[2] + s[3] * s[5], s[9]
    Ly, dt = case["lengths"][1], t2 - t1
    m0, q0, _, e0 = conserved_totals(case, a)
    m1, q1, aq1, e1 = conserved_totals(case, b)
    flux_m = np.array([s[3 * k] * s[3 * k + 1] * u_s for k in range(2)])
    return {
        "mass": np.abs(m1 - m0 - dt * Ly * flux_m) / m0,
        "energy": abs(e1 - e0 - dt * Ly * (rho_s * E_s + p_s) * u_s) / e0,
        "xmom": abs(q1[0] - q0[0] - dt * Ly * (rho_s * u_s * u_s + p_s - 1e5)) / q1[0],
        "ymom": abs(q1[1]) / aq1[0],
    }
