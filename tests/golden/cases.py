"""Case table for the golden vectors (inputs only; no solver code).

Each case is a small deck of the explicit finite-volume step: model, mesh,
order and limiters, time control and the initial domains.  Phase 0 is air
(ideal gas), phase 1 water (stiffened gas), as in the reference deck
``libTests/referenceTestCases/kapila/1D/shockTubes/interfaceWaterAir``.
"""
from __future__ import annotations

AIR = {"type": "IG", "name": "air", "gamma": 1.4, "pinf": 0.0, "cv": 800.0, "eref": 0.0, "sref": 0.0}
WATER = {"type": "SG", "name": "water", "gamma": 4.4, "pinf": 6.0e8, "cv": 1000.0, "eref": 0.0, "sref": 0.0}
HELIUM = {"type": "IG", "name": "helium", "gamma": 1.67, "pinf": 0.0, "cv": 3116.0, "eref": 0.0, "sref": 0.0}
OIL = {"type": "SG", "name": "oil", "gamma": 2.35, "pinf": 4.0e8, "cv": 1500.0, "eref": 0.0, "sref": 0.0}
# liquid water and its vapour, libEOS/SG_waterLiq.xml and IG_waterVap.xml
# (the liquid-vapour decks of libTests/referenceTestCases/eulerHomogeneous and
# kapila/AddPhysicalEffects/evap)
WATER_LIQ = {"type": "SG", "name": "waterLiq", "gamma": 3.423, "pinf": 8.99e8, "cv": 1231.2, "eref": -1.15e6,
             "sref": 0.0}
WATER_VAP = {"type": "IG", "name": "waterVap", "gamma": 1.358, "pinf": 0.0, "cv": 1247.0, "eref": 1.97e6,
             "sref": -3.56e4}


def _splitmix64(state):
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return state, z ^ (z >> 31)


def _uniforms(seed, n):
    s = seed
    out = []
    for _ in range(n):
        s, z = _splitmix64(s)
        out.append((z >> 11) * (1.0 / 9007199254740992.0))
    return out


def place_discs(seed, count, rmin, rmax, box, dx, tries=20000, log_radius=False):
    """Non-overlapping discs/spheres, centres uniform in ``box`` (list of (lo, hi)).

    Radii in cells, uniform (or log-uniform) in [rmin, rmax]; returned in metres.
    """
    import math
    placed = []
    s = seed
    for _ in range(tries):
        if len(placed) == count:
            break
        vals = []
        for _k in range(1 + len(box)):
            s, z = _splitmix64(s)
            vals.append((z >> 11) * (1.0 / 9007199254740992.0))
        if log_radius:
            r = math.exp(math.log(rmin) + vals[0] * (math.log(rmax) - math.log(rmin))) * dx
        else:
            r = (rmin + vals[0] * (rmax - rmin)) * dx
        c = [lo + r + v * (hi - lo - 2 * r) for v, (lo, hi) in zip(vals[1:], box)]
        if any(hi - lo - 2 * r <= 0 for lo, hi in box):
            continue
        ok = all(sum((a - b) ** 2 for a, b in zip(c, q)) > (r + rq + dx) ** 2 for q, rq in placed)
        if ok:
            placed.append((c, r))
    return placed


def _c1(alpha_null, t_final=240e-6, cells=1000):
    lo = {"name": "LP", "alpha": [1.0 - 1e-6, 1e-6], "rho": [50.0, 1000.0], "p": [1e5, 1e5], "u": [0.0]}
    hi = {"name": "HP", "alpha": [1e-6, 1.0 - 1e-6], "rho": [50.0, 1000.0], "p": [1e9, 1e9], "u": [0.0]}
    return {
        "model": "kapila", "alpha_null": alpha_null, "eos": [AIR, WATER],
        "cells": [cells, 1, 1], "lengths": [1.0, 0.1, 0.1], "order": 1,
        "limiters": ["none"] * 4, "cfl": 0.8, "time_mode": "time", "t_final": t_final,
        "domains": [{"type": "entireDomain", "state": "LP"},
                    {"type": "halfSpace", "state": "HP", "axe": "x", "origin": 0.7, "direction": "negative"}],
        "states": [lo, hi],
    }


def _shock_tube_1d(model, order, limiters, steps, cells=200, relax=False, alpha_null=False):
    c = _c1(alpha_null, cells=cells)
    c.update({"model": model, "order": order, "limiters": limiters, "time_mode": "iterations",
              "steps": steps, "relax_p": relax})
    if model == "multip":
        c["states"] = [{"name": "LP", "alpha": [1.0 - 1e-6, 1e-6], "rho": [50.0, 1000.0], "p": [1e5, 1e5], "u": [0.0]},
                       {"name": "HP", "alpha": [1e-6, 1.0 - 1e-6], "rho": [50.0, 1000.0], "p": [1e9, 1e9], "u": [0.0]}]
    return c


def _euler_1d(order, limiters, steps, cells=200):
    return {
        "model": "euler", "alpha_null": False, "eos": [AIR],
        "cells": [cells, 1, 1], "lengths": [1.0, 0.1, 0.1], "order": order, "limiters": limiters,
        "cfl": 0.8, "time_mode": "iterations", "steps": steps,
        "domains": [{"type": "entireDomain", "state": "LP"},
                    {"type": "halfSpace", "state": "HP", "axe": "x", "origin": 0.4, "direction": "negative"}],
        "states": [{"name": "LP", "rho": [1.0], "p": [1e5], "u": [0.0]},
                   {"name": "HP", "rho": [100.0], "p": [1e7], "u": [0.0]}],
    }


def _inj(nph, m0, axis, x_side, mix=False, rho=None, p=None):
    """bc_inj list with one injection (BoundCondInj) on side ``2*(axis-1)+x_side``
    of axis 1..3: mass flow ``m0`` per unit area into the domain of air at
    1e5 Pa (density 1.2, or 50 in the tubes' two-phase decks), or of an
    air/water mixture (alpha 0.9/0.1) when ``mix``."""
    if nph == 1:
        inj = {"m0": m0, "alpha": [1.0], "rho": [rho or 1.2], "p": [p or 1e5]}
    elif mix:
        inj = {"m0": m0, "alpha": [0.9, 0.1], "rho": [50.0, 1000.0], "p": [1e5, 1e5]}
    else:
        inj = {"m0": m0, "alpha": [1.0 - 1e-6, 1e-6], "rho": [50.0 if axis == 1 else 1.2, 1000.0], "p": [1e5, 1e5]}
    out = [None] * 6
    out[2 * (axis - 1) + x_side] = inj
    return out


def _tank(nph, p0, T0, x_side, water=False, mass=False):
    """bc_tank list with one tank (BoundCondTank) on X side ``x_side``: a
    reservoir at p0, T0 holding air (alpha 1 - 1e-6, or mass fraction 0.99 when
    ``mass``), or water when ``water``."""
    tk = {"p0": p0, "T0": T0}
    if nph > 1:
        if mass:
            tk["massFrac"] = [0.99, 0.01]
        else:
            tk["alpha"] = [1e-6, 1.0 - 1e-6] if water else [1.0 - 1e-6, 1e-6]
    out = [None] * 6
    out[x_side] = tk
    return out


def _three_phase_1d(model, order, limiters, steps, cells=150, relax=False):
    """A three-phase tube (air, water, helium): water at 1e9 Pa on the left,
    air in the middle, helium on the right, each phase at 1e-6 elsewhere."""
    f = 1e-6
    states = [{"name": "W", "alpha": [f, 1.0 - 2 * f, f], "rho": [50.0, 1000.0, 10.0], "p": [1e9] * 3, "u": [0.0]},
              {"name": "A", "alpha": [1.0 - 2 * f, f, f], "rho": [50.0, 1000.0, 10.0], "p": [1e5] * 3, "u": [0.0]},
              {"name": "H", "alpha": [f, f, 1.0 - 2 * f], "rho": [50.0, 1000.0, 10.0], "p": [1e5] * 3, "u": [0.0]}]
    return {"model": model, "alpha_null": False, "eos": [AIR, WATER, HELIUM], "cells": [cells, 1, 1],
            "lengths": [1.0, 0.1, 0.1], "order": order, "limiters": limiters, "cfl": 0.8,
            "time_mode": "iterations", "steps": steps, "relax_p": relax,
            "domains": [{"type": "entireDomain", "state": "A"},
                        {"type": "halfSpace", "state": "W", "axe": "x", "origin": 0.3, "direction": "negative"},
                        {"type": "halfSpace", "state": "H", "axe": "x", "origin": 0.7, "direction": "positive"}],
            "states": states}


def _three_phase_box(n, steps, dim, model, order, limiters, relax=False):
    """A 2-D or 3-D box of air (three phases: air, water, helium) with a
    shocked-air slab on the left, a water disc (sphere) and a helium disc
    (sphere)."""
    f = 1e-6
    p2, rho2, u2 = mach_shock_air(2.0)
    ax = [0.0] * dim
    states = [{"name": "air", "alpha": [1.0 - 2 * f, f, f], "rho": [1.2, 1000.0, 0.17], "p": [1e5] * 3, "u": ax},
              {"name": "shocked", "alpha": [1.0 - 2 * f, f, f], "rho": [rho2, 1000.0, 0.17], "p": [p2] * 3,
               "u": [u2] + ax[1:]},
              {"name": "water", "alpha": [f, 1.0 - 2 * f, f], "rho": [1.2, 1000.0, 0.17], "p": [1e5] * 3, "u": ax},
              {"name": "helium", "alpha": [f, f, 1.0 - 2 * f], "rho": [1.2, 1000.0, 0.17], "p": [1e5] * 3, "u": ax}]
    kind = "disc" if dim == 2 else "sphere"
    c1 = [0.45, 0.35] + ([0.5] if dim == 3 else [])
    c2 = [0.6, 0.7] + ([0.45] if dim == 3 else [])
    doms = [{"type": "entireDomain", "state": "air"},
            {"type": "halfSpace", "state": "shocked", "axe": "x", "origin": 0.15, "direction": "negative"},
            {"type": kind, "state": "water", "radius": 0.18, "center": c1},
            {"type": kind, "state": "helium", "radius": 0.15, "center": c2}]
    return {"model": model, "alpha_null": False, "eos": [AIR, WATER, HELIUM],
            "cells": [n, n, n if dim == 3 else 1], "lengths": [1.0, 1.0, 1.0], "order": order,
            "limiters": limiters, "cfl": 0.8, "time_mode": "iterations", "steps": steps, "relax_p": relax,
            "domains": doms, "states": states}


def _four_phase_1d(model, order, limiters, steps, cells=150, relax=False, dim2=False):
    """Four phases (air, water, helium, oil): the three-phase tube with an oil
    slab in the air; with ``dim2`` the same layers on a square grid."""
    c = _three_phase_1d(model, order, limiters, steps, cells=cells, relax=relax)
    c["eos"] = [AIR, WATER, HELIUM, OIL]
    f = 1e-6
    for s in c["states"]:
        s["alpha"] = [a * (1.0 - f) for a in s["alpha"]] + [f]
        s["rho"] = s["rho"] + [900.0]
        s["p"] = s["p"] + [s["p"][0]]
    c["states"].append({"name": "O", "alpha": [f, f, f, 1.0 - 3 * f], "rho": [50.0, 1000.0, 10.0, 900.0],
                        "p": [1e5] * 4, "u": [0.0]})
    c["domains"].append({"type": "halfSpace", "state": "O", "axe": "x", "origin": 0.45, "direction": "positive"})
    c["domains"].append({"type": "halfSpace", "state": "A", "axe": "x", "origin": 0.55, "direction": "positive"})
    c["domains"].append({"type": "halfSpace", "state": "H", "axe": "x", "origin": 0.7, "direction": "positive"})
    if dim2:
        c["cells"] = [cells, cells, 1]
        c["lengths"] = [1.0, 1.0, 1.0]
        for s in c["states"]:
            s["u"] = [0.0, 0.0]
    return c


def mach_shock_air(mach, p1=1e5, rho1=1.2, gamma=1.4):
    """Post-shock state of a normal shock moving into air at rest (Rankine-Hugoniot)."""
    import math
    m2 = mach * mach
    p2 = p1 * (1.0 + 2.0 * gamma / (gamma + 1.0) * (m2 - 1.0))
    rho2 = rho1 * (gamma + 1.0) * m2 / ((gamma - 1.0) * m2 + 2.0)
    c1 = math.sqrt(gamma * p1 / rho1)
    u2 = mach * c1 * (1.0 - rho1 / rho2)
    return p2, rho2, u2


def _shock_cylinders_2d(n=48, steps=40, ncyl=4, seed=0xEC0C0002, limiters=("vanleer",) * 4):
    p2, rho2, u2 = mach_shock_air(2.4)
    dx = 1.0 / n
    # Radii scaled to the small grid (4..8 cells) with the same placement rule.
    discs = place_discs(seed, ncyl, 4.0, 8.0, [(0.1, 0.95), (0.0, 1.0)], dx)
    doms = [{"type": "entireDomain", "state": "air"},
            {"type": "halfSpace", "state": "shocked", "axe": "x", "origin": 0.05, "direction": "negative"}]
    doms += [{"type": "disc", "state": "water", "radius": r, "center": c} for c, r in discs]
    return {
        "model": "kapila", "alpha_null": False, "eos": [AIR, WATER],
        "cells": [n, n, 1], "lengths": [1.0, 1.0, 1.0], "order": 2, "limiters": list(limiters),
        "cfl": 0.8, "time_mode": "iterations", "steps": steps, "domains": doms,
        "states": [
            {"name": "air", "alpha": [1.0 - 1e-6, 1e-6], "rho": [1.2, 1000.0], "p": [1e5, 1e5], "u": [0.0, 0.0]},
            {"name": "shocked", "alpha": [1.0 - 1e-6, 1e-6], "rho": [rho2, 1000.0], "p": [p2, p2], "u": [u2, 0.0]},
            {"name": "water", "alpha": [1e-6, 1.0 - 1e-6], "rho": [1.2, 1000.0], "p": [1e5, 1e5], "u": [0.0, 0.0]},
        ],
    }


def _bubbles_3d(n=16, steps=12, nb=3, seed=0xEC0C0003, limiters=("minmod",) * 4, model="multip", relax=True):
    dx = 1.0 / n
    spheres = place_discs(seed, nb, 2.0, 4.0, [(0.0, 1.0)] * 3, dx, log_radius=True)
    doms = [{"type": "entireDomain", "state": "water"}]
    doms += [{"type": "sphere", "state": "bubble", "radius": r, "center": c} for c, r in spheres]
    return {
        "model": model, "alpha_null": False, "eos": [AIR, WATER], "relax_p": relax,
        "cells": [n, n, n], "lengths": [1.0, 1.0, 1.0], "order": 2, "limiters": list(limiters),
        "cfl": 0.8, "time_mode": "iterations", "steps": steps, "domains": doms,
        "states": [
            {"name": "water", "alpha": [1e-6, 1.0 - 1e-6], "rho": [1.2, 1000.0], "p": [5e7, 5e7], "u": [0.0, 0.0, 0.0]},
            {"name": "bubble", "alpha": [1.0 - 1e-6, 1e-6], "rho": [1.2, 1000.0], "p": [1e5, 1e5], "u": [0.0, 0.0, 0.0]},
        ],
    }


def _water_column_2d(n, ratio, pos, t_final):
    """One 2-D case of the batched workload's distribution: a water column of
    radius pos/2 at the centre of [0,1]^2, at ratio * 1e5 Pa, in air at 1e5 Pa
    (both at rest; air density 50 as in the shock tube), Kapila, MUSCL van
    Leer, to t_final = 0.3 L / a_max.  The column's expansion puts the water
    in tension and the air at the ideal-gas pressure floor, where the
    pressure relaxation stops at its 100 iterations without converging in
    some cells (RelaxationP.cpp:92-109)."""
    hp = {"name": "HP", "alpha": [1e-6, 1.0 - 1e-6], "rho": [50.0, 1000.0], "p": [ratio * 1e5] * 2, "u": [0.0, 0.0]}
    lp = {"name": "LP", "alpha": [1.0 - 1e-6, 1e-6], "rho": [50.0, 1000.0], "p": [1e5, 1e5], "u": [0.0, 0.0]}
    return {
        "model": "kapila", "alpha_null": False, "eos": [AIR, WATER],
        "cells": [n, n, 1], "lengths": [1.0, 1.0, 1.0], "order": 2, "limiters": ["vanleer"] * 4,
        "cfl": 0.8, "time_mode": "time", "t_final": t_final,
        "domains": [{"type": "entireDomain", "state": "LP"},
                    {"type": "disc", "state": "HP", "radius": 0.5 * pos, "center": [0.5, 0.5]}],
        "states": [lp, hp],
    }


def _with_transports(case, names, values, bc_tr=None):
    """``case`` with transported scalars ``names``: ``values`` maps a state
    name to its values (0 for a state not listed, as Input.cpp:771-790 leaves
    them); ``bc_tr`` lists the values at each outflow, injection or tank side."""
    c = dict(case, transports=list(names))
    c["states"] = [dict(st, tr=list(values.get(st["name"], [0.0] * len(names)))) for st in case["states"]]
    if bc_tr is not None:
        c["bc_tr"] = bc_tr
    return c


def _interface_advection_1d():
    """The reference's transport test (libTests/referenceTestCases/kapila/1D/
    transports/interfaceWaterAir): water left of x = 0.3 and air on the right,
    both at 1e5 Pa moving at 1000 m/s, a colour of 1 in the water and 2 in the
    air, first order, CFL 0.6, to t = 500 us on 800 cells."""
    water = {"name": "water", "alpha": [1e-6, 1.0 - 1e-6], "rho": [1.0, 1000.0], "p": [1e5, 1e5], "u": [1000.0]}
    air = {"name": "air", "alpha": [1.0 - 1e-6, 1e-6], "rho": [1.0, 1000.0], "p": [1e5, 1e5], "u": [1000.0]}
    c = {
        "model": "kapila", "alpha_null": False, "eos": [AIR, WATER],
        "cells": [800, 1, 1], "lengths": [1.0, 0.1, 0.1], "order": 1, "limiters": ["none"] * 4,
        "cfl": 0.6, "time_mode": "time", "t_final": 500e-6,
        "domains": [{"type": "entireDomain", "state": "water"},
                    {"type": "halfSpace", "state": "air", "axe": "x", "origin": 0.3, "direction": "positive"}],
        "states": [air, water],
    }
    return _with_transports(c, ["color"], {"water": [1.0], "air": [2.0]})


def _thermaleq(case):
    """``case`` in the thermal-equilibrium model (ModThermalEq): each state
    keeps its volume fractions, pressure and velocity and takes the
    temperature of its air at the given density, T = p / ((gamma - 1) cv rho)."""
    c = dict(case, model="thermaleq", relax_p=False)
    c.pop("alpha_null", None)
    states = []
    for st in case["states"]:
        p = st["p"][0]
        T = p / ((AIR["gamma"] - 1.0) * AIR["cv"] * st["rho"][0])
        states.append({"name": st["name"], "alpha": list(st["alpha"]), "p": [p], "T": T, "u": list(st["u"])})
    c["states"] = states
    return c


def _eulerhom_1d(order, limiters, steps, cells=200):
    """The reference's homogeneous liquid-vapour shock tube
    (libTests/referenceTestCases/eulerHomogeneous/shockTubes): a mostly
    liquid mixture at 1e6 Pa for x < 0.3 against a mostly vapour one at 1e5 Pa."""
    return {
        "model": "eulerhomogeneous", "liq": 0, "vap": 1, "eos": [WATER_LIQ, WATER_VAP],
        "cells": [cells, 1, 1], "lengths": [1.0, 0.1, 0.1], "order": order, "limiters": limiters,
        "cfl": 0.8, "time_mode": "iterations", "steps": steps,
        "domains": [{"type": "entireDomain", "state": "LP"},
                    {"type": "halfSpace", "state": "HP", "axe": "x", "origin": 0.3, "direction": "negative"}],
        "states": [{"name": "LP", "alpha": [0.01, 0.99], "p": [1e5], "u": [0.0]},
                   {"name": "HP", "alpha": [0.99, 0.01], "p": [1e6], "u": [0.0]}],
    }


def _eulerhom_2d(order, limiters, steps, n=40):
    """A disc of mostly liquid mixture at 1e6 Pa in a mostly vapour one at 1e5 Pa."""
    c = _eulerhom_1d(order, limiters, steps)
    c.update({"cells": [n, n, 1], "lengths": [1.0, 1.0, 0.1],
              "domains": [{"type": "entireDomain", "state": "LP"},
                          {"type": "disc", "state": "HP", "radius": 0.25, "center": [0.45, 0.5]}]})
    return c


def _eulerhom_3d(steps, n=10):
    c = _eulerhom_1d(1, ["none"] * 4, steps)
    c.update({"cells": [n, n, n], "lengths": [1.0, 1.0, 1.0],
              "domains": [{"type": "entireDomain", "state": "LP"},
                          {"type": "sphere", "state": "HP", "radius": 0.3, "center": [0.45, 0.5, 0.5]}]})
    return c


def _evap_tube(order, limiters, steps, cells=200):
    """The reference's evaporation shock tube (libTests/referenceTestCases/kapila/
    AddPhysicalEffects/evap/evapShockTube): Kapila with liquid water and its
    vapour, relaxed to saturation (RelaxationPTMu) after the pressure relaxation."""
    return {
        "model": "kapila", "alpha_null": False, "eos": [WATER_LIQ, WATER_VAP], "relax_ptmu": True,
        "cells": [cells, 1, 1], "lengths": [1.0, 0.1, 0.1], "order": order, "limiters": limiters,
        "cfl": 0.8, "time_mode": "iterations", "steps": steps,
        "domains": [{"type": "entireDomain", "state": "LP"},
                    {"type": "halfSpace", "state": "HP", "axe": "x", "origin": 0.3, "direction": "negative"}],
        "states": [{"name": "LP", "alpha": [0.01, 0.99], "rho": [807.023, 0.6], "p": [1e5, 1e5], "u": [0.0]},
                   {"name": "HP", "alpha": [0.99, 0.01], "rho": [659.13, 4.894], "p": [1e6, 1e6], "u": [0.0]}],
    }


CASES = {
    # Config C1 in full: 1000 cells, first order, to t = 240 us.
    "c1_kapila_fo": _c1(False),
    "c1_kapila_fo_alphanull": _c1(True, t_final=120e-6),
    # 1-D second order, every limiter, interface limiter differing from the global one.
    "kapila_1d_o2_vanleer": _shock_tube_1d("kapila", 2, ["vanleer"] * 4, 60),
    "kapila_1d_o2_minmod_superbee": _shock_tube_1d("kapila", 2, ["minmod", "superbee", "minmod", "superbee"], 60),
    "kapila_1d_o2_vanalbada_mc": _shock_tube_1d("kapila", 2, ["vanalbada", "mc", "mc", "vanalbada"], 60),
    "kapila_1d_o2_thinc": _shock_tube_1d("kapila", 2, ["vanleer", "vanleer", "thinc", "thinc"], 60),
    "kapila_1d_o2_alphanull": _shock_tube_1d("kapila", 2, ["vanleer"] * 4, 60, alpha_null=True),
    # Pressure-disequilibrium model: first order without relaxation (reference deck), MUSCL with relaxation.
    "multip_1d_fo": _shock_tube_1d("multip", 1, ["none"] * 4, 200, cells=400),
    "multip_1d_o2_relax": _shock_tube_1d("multip", 2, ["minmod"] * 4, 60, relax=True),
    # first order with relaxation: the faces take the cells' cached states, whose
    # phase energies, sound speed, density and pressure are those before the relaxation
    "multip_1d_fo_relax": _shock_tube_1d("multip", 1, ["none"] * 4, 200, cells=400, relax=True),
    "multip_2d_fo_relax": dict(_shock_cylinders_2d(n=40, steps=60), model="multip", relax_p=True, order=1,
                               limiters=["none"] * 4),
    "multip_1d_o2_norelax": _shock_tube_1d("multip", 2, ["vanleer"] * 4, 60, relax=False),
    # Single-phase Euler.
    "euler_1d_fo": _euler_1d(1, ["none"] * 4, 100),
    "euler_1d_o2": _euler_1d(2, ["vanleer"] * 4, 100),
    # Multi-D: config C2 physics (shock on water cylinders) and C3 physics (bubbles, MultiP + relaxation).
    "c2_kapila_2d_o2": _shock_cylinders_2d(),
    "c2_kapila_2d_fo": dict(_shock_cylinders_2d(n=40, steps=40), order=1, limiters=["none"] * 4),
    "c3_multip_3d_o2": _bubbles_3d(),
    # Case 1680 of the batched workload (seed 0xEC0C0004): the relaxation's
    # iteration limit is reached in 124 cell updates.
    "kapila_2d_o2_relax_limit": _water_column_2d(64, 2679.014153940562, 0.35536181384193916, 0.0001535181427478315),
    "kapila_3d_o2": _bubbles_3d(n=12, steps=10, model="kapila", relax=False, limiters=("vanleer",) * 4),
    # Wall boundaries (first order: BoundCondWall, also the reference's first-order SYMMETRY):
    # the C1 tube closed at both ends, run until the air shock has reflected off x = 1;
    # MultiP and Euler tubes with one wall; the 2-D cylinders in a closed box.
    "c1_kapila_fo_wall": dict(_c1(False, t_final=600e-6, cells=400), bc=["wall"] * 6),
    "multip_1d_fo_wall": dict(_shock_tube_1d("multip", 1, ["none"] * 4, 300, cells=200), bc=["wall", "abs"] + ["abs"] * 4),
    "euler_1d_fo_wall": dict(_euler_1d(1, ["none"] * 4, 300), bc=["abs", "wall"] + ["abs"] * 4),
    "kapila_2d_fo_wall": dict(_shock_cylinders_2d(n=40, steps=80), order=1, limiters=["none"] * 4, bc=["wall"] * 6),
    # Walls at second order (BoundCondWallO2: mirrored normal-velocity slopes, the cell
    # extrapolated to the wall), with the interface limiters and THINC next to a wall.
    "kapila_1d_o2_wall": dict(_shock_tube_1d("kapila", 2, ["vanleer"] * 4, 400, cells=200), bc=["wall"] * 6),
    # (interface limiters next to walls: minmod / van Leer; the minmod / superbee pair is
    # chaotic over this length -- a 2e-16 change of the initial pressure moves the
    # result by 6e-2 in the oracle itself -- and so cannot pin 1e-12 parity)
    "kapila_1d_o2_minmod_vanleer_wall": dict(
        _shock_tube_1d("kapila", 2, ["minmod", "vanleer", "minmod", "vanleer"], 200, cells=100), bc=["wall"] * 6),
    "kapila_1d_o2_thinc_wall": dict(
        _shock_tube_1d("kapila", 2, ["vanleer", "vanleer", "thinc", "thinc"], 400, cells=200), bc=["wall"] * 6),
    "multip_1d_o2_relax_wall": dict(_shock_tube_1d("multip", 2, ["minmod"] * 4, 160, cells=100, relax=True),
                                    bc=["wall", "abs"] + ["abs"] * 4),
    "euler_1d_o2_wall": dict(_euler_1d(2, ["vanleer"] * 4, 300), bc=["abs", "wall"] + ["abs"] * 4),
    "kapila_2d_o2_wall": dict(_shock_cylinders_2d(n=40, steps=50), bc=["wall"] * 6),
    # Outflow boundaries (BoundCondOutflow, imposed exit pressure; no second-order
    # variant): the tubes exit through x+ into a pressure below the air's, and the
    # Euler tube through x- against the expansion.
    "kapila_1d_fo_outflow": dict(_shock_tube_1d("kapila", 1, ["none"] * 4, 300, cells=200),
                                 bc=["abs", "outflow"] + ["abs"] * 4, bc_p0=[0.0, 5e4] + [0.0] * 4),
    "kapila_1d_o2_outflow": dict(_shock_tube_1d("kapila", 2, ["vanleer"] * 4, 300, cells=200),
                                 bc=["outflow", "outflow"] + ["abs"] * 4, bc_p0=[2e9, 5e4] + [0.0] * 4),
    "multip_1d_o2_relax_outflow": dict(_shock_tube_1d("multip", 2, ["minmod"] * 4, 160, cells=100, relax=True),
                                       bc=["abs", "outflow"] + ["abs"] * 4, bc_p0=[0.0, 5e4] + [0.0] * 4),
    "euler_1d_fo_outflow": dict(_euler_1d(1, ["none"] * 4, 300), bc=["outflow", "outflow"] + ["abs"] * 4,
                                bc_p0=[2e7, 5e4] + [0.0] * 4),
    "kapila_2d_fo_outflow": dict(_shock_cylinders_2d(n=40, steps=80), order=1, limiters=["none"] * 4,
                                 bc=["abs", "outflow", "wall", "wall", "abs", "abs"], bc_p0=[0.0, 8e4] + [0.0] * 4),
    # Injection boundaries (BoundCondInj, imposed mass flow of a given fluid; no
    # second-order variant): air, or an air/water mixture, enters the tubes'
    # low-pressure end, and air enters the 2-D box through its bottom face.
    "kapila_1d_fo_injection": dict(_shock_tube_1d("kapila", 1, ["none"] * 4, 300, cells=200),
                                   bc=["abs", "injection"] + ["abs"] * 4, bc_inj=_inj(2, 500.0, 1, x_side=1)),
    "kapila_1d_o2_injection": dict(_shock_tube_1d("kapila", 2, ["vanleer"] * 4, 300, cells=200),
                                   bc=["wall", "injection"] + ["abs"] * 4, bc_inj=_inj(2, 2000.0, 1, x_side=1, mix=True)),
    "multip_1d_o2_relax_injection": dict(_shock_tube_1d("multip", 2, ["minmod"] * 4, 160, cells=100, relax=True),
                                         bc=["abs", "injection"] + ["abs"] * 4, bc_inj=_inj(2, 800.0, 1, x_side=1)),
    "euler_1d_fo_injection": dict(_euler_1d(1, ["none"] * 4, 300), bc=["abs", "injection"] + ["abs"] * 4,
                                  bc_inj=_inj(1, 150.0, 1, x_side=1)),
    "euler_1d_o2_injection": dict(_euler_1d(2, ["vanleer"] * 4, 300), bc=["injection", "outflow"] + ["abs"] * 4,
                                  bc_p0=[0.0, 5e4] + [0.0] * 4, bc_inj=_inj(1, 3000.0, 1, x_side=0, rho=80.0, p=1e7)),
    # Three phases (air, water, helium)
    "kapila_1d_fo_3ph": _three_phase_1d("kapila", 1, ["none"] * 4, 200),
    "kapila_1d_o2_3ph": _three_phase_1d("kapila", 2, ["vanleer", "minmod", "vanleer", "superbee"], 200),
    "multip_1d_o2_relax_3ph": _three_phase_1d("multip", 2, ["minmod"] * 4, 150, cells=100, relax=True),
    "multip_1d_o2_norelax_3ph": _three_phase_1d("multip", 2, ["vanleer"] * 4, 150, cells=100),
    "kapila_2d_o2_3ph": _three_phase_box(32, 30, 2, "kapila", 2, ["vanleer"] * 4),
    "kapila_1d_o2_4ph": _four_phase_1d("kapila", 2, ["vanleer"] * 4, 200),
    "multip_2d_o2_relax_4ph": _four_phase_1d("multip", 2, ["minmod"] * 4, 30, cells=24, relax=True, dim2=True),
    "multip_3d_o2_relax_3ph": _three_phase_box(10, 8, 3, "multip", 2, ["minmod"] * 4, relax=True),
    # Tank boundaries (BoundCondTank, a reservoir at rest at p0, T0; no
    # second-order variant): air flows in from tanks above the tubes' air
    # pressure, and the tubes' high-pressure water flows out into tanks below
    # it (the subsonic-outflow branch); one MultiP tank is given by mass fractions.
    "kapila_1d_fo_tank": dict(_shock_tube_1d("kapila", 1, ["none"] * 4, 300, cells=200),
                              bc=["abs", "tank"] + ["abs"] * 4, bc_tank=_tank(2, 3e5, 300.0, 1)),
    "kapila_1d_o2_tank": dict(_shock_tube_1d("kapila", 2, ["vanleer"] * 4, 300, cells=200), bc=["tank", "tank"] + ["abs"] * 4,
                              bc_tank=[_tank(2, 5e8, 300.0, 0, water=True)[0], _tank(2, 3e5, 300.0, 1)[1]] + [None] * 4),
    "multip_1d_o2_relax_tank": dict(_shock_tube_1d("multip", 2, ["minmod"] * 4, 160, cells=100, relax=True),
                                    bc=["abs", "tank"] + ["abs"] * 4, bc_tank=_tank(2, 4e5, 350.0, 1, mass=True)),
    "euler_1d_fo_tank": dict(_euler_1d(1, ["none"] * 4, 300), bc=["abs", "tank"] + ["abs"] * 4,
                             bc_tank=_tank(1, 3e5, 300.0, 1)),
    "euler_1d_o2_tank": dict(_euler_1d(2, ["vanleer"] * 4, 300), bc=["tank", "tank"] + ["abs"] * 4,
                             bc_tank=[_tank(1, 5e6, 300.0, 0)[0], _tank(1, 2e5, 300.0, 1)[1]] + [None] * 4),
    "kapila_2d_fo_tank": dict(_shock_cylinders_2d(n=40, steps=80), order=1, limiters=["none"] * 4,
                              bc=["abs", "outflow", "tank", "wall", "abs", "abs"],
                              bc_p0=[0.0, 8e4] + [0.0] * 4, bc_tank=[None, None] + _tank(2, 3e5, 300.0, 0)[:1] + [None] * 3),
    "kapila_2d_fo_injection": dict(_shock_cylinders_2d(n=40, steps=80), order=1, limiters=["none"] * 4,
                                   bc=["abs", "outflow", "injection", "wall", "abs", "abs"],
                                   bc_p0=[0.0, 8e4] + [0.0] * 4, bc_inj=_inj(2, 300.0, 2, x_side=0)),
    # Transported scalars (Transport, src/Transport/Transport.cpp): the
    # reference's own interface-advection test, then MUSCL with the interface
    # volume-fraction limiter (THINC, minmod slopes for the scalars) next to a
    # different global one, and every boundary type carrying scalars in and out
    "kapila_1d_fo_transport": _interface_advection_1d(),
    "kapila_1d_o2_transport_thinc": _with_transports(
        _shock_tube_1d("kapila", 2, ["vanleer", "superbee", "mc", "thinc"], 120), ["color", "dye"],
        {"LP": [1.0, 0.25], "HP": [0.0, -3.0]}),
    "kapila_1d_o2_transport_wall_tank": _with_transports(
        dict(_shock_tube_1d("kapila", 2, ["minmod", "vanleer", "minmod", "vanleer"], 300, cells=200),
             bc=["wall", "tank"] + ["abs"] * 4, bc_tank=_tank(2, 3e5, 300.0, 1)),
        ["color"], {"LP": [1.0], "HP": [0.0]}, bc_tr=[None, [5.0]] + [None] * 4),
    "multip_1d_o2_relax_transport_inj_out": _with_transports(
        dict(_shock_tube_1d("multip", 2, ["minmod"] * 4, 160, cells=100, relax=True),
             bc=["outflow", "injection"] + ["abs"] * 4, bc_p0=[2e9, 0.0] + [0.0] * 4, bc_inj=_inj(2, 800.0, 1, x_side=1)),
        ["color", "dye"], {"LP": [1.0, 2.0], "HP": [0.0, 1.0]}, bc_tr=[[0.5, 7.0], [3.0, -1.0]] + [None] * 4),
    "kapila_1d_fo_transport_outflow": _with_transports(
        dict(_shock_tube_1d("kapila", 1, ["none"] * 4, 300, cells=200),
             bc=["outflow", "outflow"] + ["abs"] * 4, bc_p0=[2e9, 5e4] + [0.0] * 4),
        ["color"], {"LP": [1.0], "HP": [0.0]}, bc_tr=[[4.0], [9.0]] + [None] * 4),
    "kapila_2d_o2_transport": _with_transports(
        dict(_shock_cylinders_2d(n=40, steps=50), bc=["abs", "outflow", "wall", "wall", "abs", "abs"],
             bc_p0=[0.0, 8e4] + [0.0] * 4),
        ["color", "dye"], {"water": [1.0, 0.0], "shocked": [0.0, 1.0]}, bc_tr=[None, [0.0, 0.0]] + [None] * 4),
    "multip_2d_fo_relax_transport": _with_transports(
        dict(_shock_cylinders_2d(n=40, steps=60), model="multip", relax_p=True, order=1, limiters=["none"] * 4,
             bc=["abs", "outflow", "injection", "wall", "abs", "abs"], bc_p0=[0.0, 8e4] + [0.0] * 4,
             bc_inj=_inj(2, 300.0, 2, x_side=0)),
        ["color"], {"water": [1.0], "shocked": [0.5]}, bc_tr=[None, [0.0], [2.0]] + [None] * 3),
    "multip_3d_o2_transport": _with_transports(_bubbles_3d(n=12, steps=10), ["color"], {"bubble": [1.0]}),
    # Thermal-equilibrium model (ModThermalEq, FluxThermalEq, MixThermalEq):
    # two phases at one pressure and temperature, the mixture HLLC, the
    # closed-form pressure of buildPrim; first order and MUSCL (THINC with an
    # interface limiter), walls and outflows, 1-D/2-D/3-D
    "thermaleq_1d_fo": _thermaleq(_shock_tube_1d("kapila", 1, ["none"] * 4, 200, cells=200)),
    "thermaleq_1d_o2_vanleer": _thermaleq(_shock_tube_1d("kapila", 2, ["vanleer"] * 4, 120)),
    "thermaleq_1d_o2_thinc": _thermaleq(_shock_tube_1d("kapila", 2, ["minmod", "vanleer", "minmod", "thinc"], 120)),
    "thermaleq_1d_o2_wall_outflow": _thermaleq(dict(_shock_tube_1d("kapila", 2, ["minmod", "vanleer", "minmod", "vanleer"], 300),
                                                    bc=["wall", "outflow"] + ["abs"] * 4, bc_p0=[0.0, 5e4] + [0.0] * 4)),
    "thermaleq_1d_fo_outflow": _thermaleq(dict(_shock_tube_1d("kapila", 1, ["none"] * 4, 300),
                                               bc=["outflow", "outflow"] + ["abs"] * 4, bc_p0=[2e9, 5e4] + [0.0] * 4)),
    "thermaleq_2d_fo_wall": _thermaleq(dict(_shock_cylinders_2d(n=40, steps=60), order=1, limiters=["none"] * 4,
                                            bc=["wall"] * 6)),
    "thermaleq_2d_o2": _thermaleq(dict(_shock_cylinders_2d(n=40, steps=40), bc=["abs", "outflow", "wall", "wall", "abs", "abs"],
                                       bc_p0=[0.0, 8e4] + [0.0] * 4)),
    "thermaleq_3d_o2": _thermaleq(_bubbles_3d(n=12, steps=10, model="kapila", relax=False, limiters=("vanleer",) * 4)),
    "thermaleq_2d_o2_abs": _thermaleq(_shock_cylinders_2d(n=40, steps=30, limiters=("minmod",) * 4)),
    # Pressure-temperature relaxation (RelaxationPT) after the pressure one:
    # Kapila (which relaxes P itself) and MultiP with P then PT, or PT alone
    "kapila_1d_fo_relax_pt": dict(_shock_tube_1d("kapila", 1, ["none"] * 4, 200), relax_pt=True),
    "kapila_1d_o2_relax_pt": dict(_shock_tube_1d("kapila", 2, ["vanleer"] * 4, 120), relax_pt=True),
    "multip_1d_o2_relax_p_pt": dict(_shock_tube_1d("multip", 2, ["minmod"] * 4, 120, relax=True), relax_pt=True),
    "multip_1d_o2_pt": dict(_shock_tube_1d("multip", 2, ["vanleer"] * 4, 120), relax_pt=True),
    "multip_2d_fo_relax_p_pt": dict(_shock_cylinders_2d(n=40, steps=50), model="multip", relax_p=True, relax_pt=True,
                                    order=1, limiters=["none"] * 4),
    "kapila_2d_o2_relax_pt": dict(_shock_cylinders_2d(n=40, steps=40), relax_pt=True,
                                  bc=["abs", "outflow", "wall", "wall", "abs", "abs"], bc_p0=[0.0, 8e4] + [0.0] * 4),
    # Homogeneous liquid-vapour model (ModEulerHomogeneous), first order (the
    # reference's MUSCL form starts its first predictor's Newton solve from an
    # unset pressure and prints NaN): the reference's shock tube, iteration and
    # time controlled, a 2-D disc and a 3-D sphere
    "eulerhom_1d_fo": _eulerhom_1d(1, ["none"] * 4, 200),
    "eulerhom_1d_fo_time": dict(_eulerhom_1d(1, ["none"] * 4, 0, cells=400), time_mode="time", t_final=1.2e-3),
    "eulerhom_2d_fo": _eulerhom_2d(1, ["none"] * 4, 60),
    "eulerhom_3d_fo": _eulerhom_3d(12),
    # Relaxation to saturation (RelaxationPTMu) after Kapila's pressure relaxation:
    # the reference's evaporation shock tube at first order and MUSCL
    "kapila_1d_fo_ptmu": _evap_tube(1, ["none"] * 4, 200),
    "kapila_1d_o2_ptmu": _evap_tube(2, ["vanleer"] * 4, 150),
    "kapila_2d_o2_ptmu": dict(_evap_tube(2, ["minmod"] * 4, 30), cells=[32, 32, 1], lengths=[1.0, 1.0, 0.1],
                              domains=[{"type": "entireDomain", "state": "LP"},
                                       {"type": "disc", "state": "HP", "radius": 0.25, "center": [0.45, 0.5]}]),
}


# Runs the reference stops (Errors::errorMessage, exit) because an injection
# Newton solve does not converge in 50 iterations: a mass flow no left wave
# can carry (ModKapila.cpp:283, ModEuler::solveRiemannInflow).  Output at
# every iteration, so the last infoCalcul.out row is the last completed step.
STOP_CASES = {
    "kapila_1d_fo_injection_stop": dict(CASES["kapila_1d_fo_injection"], bc_inj=_inj(2, 1e6, 1, x_side=1), iter_freq=1),
    "euler_1d_fo_injection_stop": dict(CASES["euler_1d_fo_injection"], bc_inj=_inj(1, 1e4, 1, x_side=1), iter_freq=1),
}
