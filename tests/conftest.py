"""Shared pytest configuration: the ``gpu`` marker and golden-case helpers."""
import json
import sys
from pathlib import Path

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
    else:
        out["rho"] = sum(state[3 * k] * state[3 * k + 1] for k in range(nph))
        out["p"] = sum(state[3 * k] * state[3 * k + 2] for k in range(nph))
        for k in range(nph):
            out[f"alpha{k}"] = state[3 * k]
    out["velocity"] = np.concatenate([state[3 * nph + d] for d in range(dim)])
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
    return out
