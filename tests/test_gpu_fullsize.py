"""Size-independent properties of the HIP path at the full sizes of the 2-D
shock/cylinders and 3-D bubble workloads, where the oracle is too slow to
march the same steps in a test.

The update is conservative (`src/Run.cpp:356-412`: each face flux leaves one
cell and enters its neighbour), so per-phase mass and mixture total energy
change only by what crosses the boundary:

* 2-D, 4096^2, Kapila MUSCL van Leer: a planar Mach 2.4 shock enters through
  the left boundary.  The post-shock flow is supersonic (u - c > 0), so the
  left boundary column keeps its initial state and its transmissive face
  carries the physical flux F(U_shocked) exactly; the right boundary is air
  at rest (no mass or energy flux, pressure 1e5 on the momentum).  After 10
  steps nothing else has reached a boundary, and no velocity in y has
  appeared (the flow is still planar).
* 3-D, 384^3, MultiP with relaxation, MUSCL minmod: the 64 bubbles are
  placed at least 0.1 from every face, so after 5 steps every boundary cell is
  still water at rest: nothing crosses the boundary, and total momentum stays
  zero to round-off.

The tolerances are round-off bounds: 1e-12 relative for mass and energy
(the oracle gives ~1e-16 on the same problems at 256^2 and 96^3,
``tests/test_oracle_kat.py``), 1e-11 for the
x-momentum balance (a sum with cancellations), and 1e-10 of sum rho|u| for
momenta that must stay zero.  The FP32-storage form is checked at the same
sizes against the FP64 path (relative L1 <= 1e-5 in mixture density,
velocity, mixture pressure and volume fractions) and for the same balances
to 1e-8 (1e-7 for x-momentum): each step rounds the changed cells' alpha_k,
rho_k, p_k, u and reduced energy to 24 bits, and uniform regions do not
change and do not round again.
"""
import numpy as np
import pytest

from conftest import conserved_totals, parity_quantities, rel_l1, shock_inflow_errors

pytestmark = pytest.mark.gpu

TOL_CONS = 1e-12
TOL_XMOM = 1e-11
TOL_ZERO = 1e-10
TOL_FP32S = 1e-5
TOL_CONS_FP32S = 1e-8


def _march(case, st, steps):
    """One step (the reference's first step, dt = 1e-15, which also yields E),
    then ``steps`` more: (state after the first step, t there, final state, final t)."""
    from ecogen_amd import Context
    with Context(case) as ctx:
        ctx.set_state(st)
        s1 = ctx.run(nsteps=1)
        a = ctx.get_state()
        s2 = ctx.run(nsteps=steps)
        assert s2.steps == steps + 1
        assert s2.flags & 2 == 0
        b = ctx.get_state()
    return a, s1.t, b, s2.t


@pytest.fixture(scope="module")
def c2():
    from ecogen_amd import workloads as W
    case, st = W.shock_cylinders_2d(n=4096, steps=10)
    print("case shock_cylinders_2d 4096^2 fp64, 1+10 steps", flush=True)
    return case, st, _march(case, st, 10)


def _check_c2_balance(case, a, t1, b, t2, tol_mass, tol_energy, tol_xmom):
    err = shock_inflow_errors(case, a, t1, b, t2)
    print("  " + ", ".join(f"{k} {v}" for k, v in err.items()), flush=True)
    assert np.all(err["mass"] <= tol_mass), err
    assert err["energy"] <= tol_energy, err
    assert err["xmom"] <= tol_xmom, err
    assert err["ymom"] <= TOL_ZERO, err


def test_c2_full_size_conservation(c2):
    case, _, (a, t1, b, t2) = c2
    _check_c2_balance(case, a, t1, b, t2, TOL_CONS, TOL_CONS, TOL_XMOM)


def test_c2_full_size_fp32s(c2):
    """FP32 storage at 4096^2: within 1e-5 of the FP64 path, and mass-balanced."""
    case, st, (_, _, b64, t64) = c2
    c32 = dict(case, precision="fp32s")
    a, t1, b, t2 = _march(c32, st, 10)
    assert t2 == pytest.approx(t64, rel=1e-5)
    got, want = parity_quantities(case, b), parity_quantities(case, b64)
    for nm in got:
        e = rel_l1(got[nm], want[nm])
        print(f"  {nm}: rel L1 vs FP64 {e:.2e}", flush=True)
        assert e <= TOL_FP32S, f"{nm}: {e:.3e}"
    _check_c2_balance(c32, a, t1, b, t2, TOL_CONS_FP32S, TOL_CONS_FP32S, 10 * TOL_CONS_FP32S)


def test_c3_full_size_conservation():
    from ecogen_amd import workloads as W
    case, st = W.bubbles_3d(n=384, steps=5, box=(0.1, 0.9))
    print("case bubbles_3d 384^3 fp64 (bubbles 0.1 from the faces), 1+5 steps", flush=True)
    a, _, b, _ = _march(case, st, 5)
    del st
    m0, _, _, e0 = conserved_totals(case, a)
    m1, q1, aq1, e1 = conserved_totals(case, b)
    err_m = np.abs(m1 - m0) / m0
    err_e = abs(e1 - e0) / abs(e0)
    print(f"  mass {err_m}, energy {err_e:.2e}, momentum {q1} of {aq1}", flush=True)
    assert np.all(err_m <= TOL_CONS), err_m
    assert err_e <= TOL_CONS, err_e
    assert np.all(np.abs(q1) <= TOL_ZERO * aq1)


def test_c2_full_size_injection_mass_inflow():
    """4096^2, first order: air enters the closed box through its bottom face
    (BoundCondInj, m0 = 300 kg/s/m^2).  The injection face's mass flux is m0
    up to round-off and every other side is a wall, so phase k's mass grows by
    Y_k0 m0 S t exactly (S = 1 m^2, t the time marched)."""
    from ecogen_amd import Context
    from ecogen_amd import workloads as W
    case, st = W.shock_cylinders_2d(n=4096, steps=10)
    inj = {"m0": 300.0, "alpha": [1.0 - 1e-6, 1e-6], "rho": [1.2, 1000.0], "p": [1e5, 1e5]}
    case = dict(case, order=1, limiters=["none"] * 4, bc=["wall", "wall", "injection", "wall", "abs", "abs"],
                bc_inj=[None, None, inj, None, None, None])
    print("case shock_cylinders_2d 4096^2 fp64 first order, injection at y-, walls elsewhere, 10 steps", flush=True)
    vol = (1.0 / 4096) ** 2
    m_start = [float((st[3 * k] * st[3 * k + 1]).sum()) * vol for k in range(2)]
    with Context(case) as ctx:
        ctx.set_state(st)
        s = ctx.run(nsteps=10)
        assert s.steps == 10 and s.flags == 0
        out = ctx.get_state()
    rho0 = inj["alpha"][0] * inj["rho"][0] + inj["alpha"][1] * inj["rho"][1]
    for k in range(2):
        m_end = float((out[3 * k] * out[3 * k + 1]).sum()) * vol
        want = inj["alpha"][k] * inj["rho"][k] / rho0 * inj["m0"] * 1.0 * s.t
        print(f"  phase {k}: mass gained {m_end - m_start[k]:.6e}, m0 Y S t {want:.6e}", flush=True)
        assert abs((m_end - m_start[k]) - want) <= 1e-12 * m_end + 1e-9 * want


def test_c2_full_size_tank_equilibrium():
    """4096^2, first order, a tank (BoundCondTank) at p0 = 2e5 on every side
    of a box of air/water mixture at rest at 2e5: the tanks meet the subsonic
    outflow branch with p* = p0 and a contact speed of round-off size, so the
    box stays at rest and at p0 (the oracle shows the same on a tube)."""
    from cases import AIR, WATER
    from ecogen_amd import Context
    n = 4096
    tank = {"p0": 2e5, "T0": 300.0, "alpha": [0.7, 0.3]}
    case = {"model": "kapila", "alpha_null": False, "eos": [AIR, WATER], "cells": [n, n, 1],
            "lengths": [1.0, 1.0, 1.0], "order": 1, "limiters": ["none"] * 4, "cfl": 0.8,
            "time_mode": "iterations", "steps": 10, "bc": ["tank"] * 4 + ["abs"] * 2,
            "bc_tank": [tank] * 4 + [None] * 2}
    st = np.zeros((10, n * n))
    st[0], st[1], st[3], st[4] = 0.7, 2.3, 0.3, 1000.0
    st[2] = st[5] = 2e5
    print("case 4096^2 mixture at rest, tanks at p0 on all four sides, fp64 first order, 10 steps", flush=True)
    with Context(case) as ctx:
        ctx.set_state(st)
        s = ctx.run(nsteps=10)
        assert s.steps == 10 and s.flags == 0
        out = ctx.get_state()
    umax = float(np.max(np.abs(out[6:8])))
    dp = float(np.max(np.abs(out[2] - 2e5)))
    print(f"  max |u| {umax:.2e}, max |p - p0| {dp:.2e}", flush=True)
    assert umax < 1e-9 and dp < 1e-9 * 2e5
