"""Synthetic initial states of the benchmark configurations.

Initial conditions follow the reference's geometric-domain semantics
(GeometricalDomain::fillIn, src/Geometries/GeometricalDomain.cpp:70-88: later
domains override earlier ones; GDHalfSpace::belong ``x <= origin`` for
direction negative, GDHalfSpace.cpp:74-96; GDDisc::belong / GDSphere::belong
``sum of squared offsets <= r*r``, GDDisc.cpp:85-107, GDSphere.cpp:71-79) on
the reference cell centres ``(i + 0.5) * L / n``.  Placement of cylinders and
bubbles is seeded (splitmix64) and non-overlapping.

State layout: ecogen.h host layout, field-major (alpha, rho, p) per phase,
then u, v, w (E is computed on the device).
"""
from __future__ import annotations

import math

import numpy as np

AIR = {"type": "IG", "name": "air", "gamma": 1.4, "pinf": 0.0, "cv": 800.0, "eref": 0.0}
WATER = {"type": "SG", "name": "water", "gamma": 4.4, "pinf": 6.0e8, "cv": 1000.0, "eref": 0.0}
ALPHA_FLOOR = 1e-6


class SplitMix64:
    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)


def place(seed, count, rmin_cells, rmax_cells, box, dx, log_radius=False, tries=200000):
    """Non-overlapping discs/spheres: centres uniform inside ``box`` (list of (lo, hi)) keeping
    the body inside, radii uniform (or log-uniform) in cells. Returns [(centre, radius_m)]."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(tries):
        if len(out) == count:
            break
        u = rng.uniform()
        if log_radius:
            r = math.exp(math.log(rmin_cells) + u * (math.log(rmax_cells) - math.log(rmin_cells))) * dx
        else:
            r = (rmin_cells + u * (rmax_cells - rmin_cells)) * dx
        c = [lo + r + rng.uniform() * (hi - lo - 2 * r) for lo, hi in box]
        if all(sum((a - b) ** 2 for a, b in zip(c, q)) > (r + rq + dx) ** 2 for q, rq in out):
            out.append((c, r))
    if len(out) < count:
        raise RuntimeError(f"could only place {len(out)} of {count} bodies")
    return out


def mach_shock_air(mach, p1=1e5, rho1=1.2, gamma=1.4):
    """Post-shock state behind a normal shock running into air at rest (Rankine-Hugoniot)."""
    m2 = mach * mach
    p2 = p1 * (1.0 + 2.0 * gamma / (gamma + 1.0) * (m2 - 1.0))
    rho2 = rho1 * (gamma + 1.0) * m2 / ((gamma - 1.0) * m2 + 2.0)
    c1 = math.sqrt(gamma * p1 / rho1)
    return p2, rho2, mach * c1 * (1.0 - rho1 / rho2)


def _centres(n, L):
    return [(np.arange(n[d]) + 0.5) * (L[d] / n[d]) for d in range(3)]


def _vals(st, nph=2):
    out = []
    for k in range(nph):
        out += [st["alpha"][k], st["rho"][k], st["p"][k]]
    return out + list(st["u"])


def _fill(state, mask_idx, st, nph=2):
    for k in range(nph):
        state[3 * k][mask_idx] = st["alpha"][k]
        state[3 * k + 1][mask_idx] = st["rho"][k]
        state[3 * k + 2][mask_idx] = st["p"][k]
    for d, u in enumerate(st["u"]):
        state[3 * nph + d][mask_idx] = u


def cylinder_bodies(n=4096, ncyl=32, seed=0xEC0C0002, rmin=20.0, rmax=80.0):
    """Water cylinders of shock_cylinders_2d: ((cx, cy), r), non-overlapping,
    centres in x in [0.1, 0.95], radii uniform in [rmin, rmax] cells."""
    return place(seed, ncyl, rmin, rmax, [(0.1, 0.95), (0.0, 1.0)], 1.0 / n)


def shock_cylinders_2d_states():
    """(air, shocked air, water) states of shock_cylinders_2d."""
    p2, rho2, u2 = mach_shock_air(2.4)
    air = {"name": "air", "alpha": [1 - ALPHA_FLOOR, ALPHA_FLOOR], "rho": [1.2, 1000.0], "p": [1e5, 1e5],
           "u": [0.0, 0.0, 0.0]}
    shocked = {"name": "shocked", "alpha": [1 - ALPHA_FLOOR, ALPHA_FLOOR], "rho": [rho2, 1000.0], "p": [p2, p2],
               "u": [u2, 0.0, 0.0]}
    water = {"name": "water", "alpha": [ALPHA_FLOOR, 1 - ALPHA_FLOOR], "rho": [1.2, 1000.0], "p": [1e5, 1e5],
             "u": [0.0, 0.0, 0.0]}
    return air, shocked, water


def shock_cylinders_2d(n=4096, ncyl=32, seed=0xEC0C0002, rmin=20.0, rmax=80.0, limiter="vanleer", steps=1000):
    """Config C2: a planar Mach 2.4 air shock (x < 0.05) hits ``ncyl`` water cylinders."""
    case = {
        "model": "kapila", "alpha_null": False, "eos": [AIR, WATER],
        "cells": [n, n, 1], "lengths": [1.0, 1.0, 1.0], "order": 2, "limiters": [limiter] * 4,
        "cfl": 0.8, "time_mode": "iterations", "steps": steps,
    }
    L = case["lengths"]
    xc, yc, _ = _centres(case["cells"], L)
    ncell = n * n
    st = np.empty((9, ncell))
    air, shocked, water = shock_cylinders_2d_states()
    _fill(st, slice(None), air)
    view = st.reshape(9, n, n)  # [field, j, i]
    sx = xc <= 0.05
    for f, v in enumerate(_vals(shocked)):
        view[f][:, sx] = v
    dx = L[0] / n
    for (cx, cy), r in cylinder_bodies(n, ncyl, seed, rmin, rmax):
        i0, i1 = max(0, int((cx - r) / dx) - 1), min(n, int((cx + r) / dx) + 2)
        j0, j1 = max(0, int((cy - r) / dx) - 1), min(n, int((cy + r) / dx) + 2)
        X = xc[i0:i1][None, :]
        Y = yc[j0:j1][:, None]
        inside = (X - cx) ** 2 + (Y - cy) ** 2 <= r * r
        for f, v in enumerate(_vals(water)):
            blk = view[f][j0:j1, i0:i1]
            blk[inside] = v
    return case, st


def bubbles_3d(n=384, nb=64, seed=0xEC0C0003, rmin=4.0, rmax=24.0, limiter="minmod", steps=500, precision="fp64",
               box=(0.0, 1.0)):
    """Config C3: ``nb`` air bubbles in water at 5e7 Pa, pressure-disequilibrium model with relaxation.
    The bubbles lie inside ``box`` along every axis (the whole domain by default)."""
    case = {
        "model": "multip", "alpha_null": False, "eos": [AIR, WATER], "relax_p": True,
        "cells": [n, n, n], "lengths": [1.0, 1.0, 1.0], "order": 2, "limiters": [limiter] * 4,
        "cfl": 0.8, "time_mode": "iterations", "steps": steps, "precision": precision,
    }
    L = case["lengths"]
    xc, yc, zc = _centres(case["cells"], L)
    st = np.empty((9, n * n * n))
    water = {"alpha": [ALPHA_FLOOR, 1 - ALPHA_FLOOR], "rho": [1.2, 1000.0], "p": [5e7, 5e7], "u": [0.0, 0.0, 0.0]}
    bubble = [1 - ALPHA_FLOOR, 1.2, 1e5, ALPHA_FLOOR, 1000.0, 1e5, 0.0, 0.0, 0.0]
    _fill(st, slice(None), water)
    view = st.reshape(9, n, n, n)  # [field, k, j, i]
    dx = L[0] / n
    for (cx, cy, cz), r in place(seed, nb, rmin, rmax, [tuple(box)] * 3, dx, log_radius=True):
        lo = [max(0, int((c - r) / dx) - 1) for c in (cx, cy, cz)]
        hi = [min(n, int((c + r) / dx) + 2) for c in (cx, cy, cz)]
        X = xc[lo[0]:hi[0]][None, None, :]
        Y = yc[lo[1]:hi[1]][None, :, None]
        Z = zc[lo[2]:hi[2]][:, None, None]
        inside = (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2 <= r * r
        for f, v in enumerate(bubble):
            blk = view[f][lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
            blk[inside] = v
    return case, st


def shock_tube_1d(n=1000):
    """Config C1: water (x <= 0.7, 1e9 Pa) / air (1e5 Pa) shock tube, first order, to 240 us."""
    case = {
        "model": "kapila", "alpha_null": False, "eos": [AIR, WATER],
        "cells": [n, 1, 1], "lengths": [1.0, 0.1, 0.1], "order": 1, "limiters": ["none"] * 4,
        "cfl": 0.8, "time_mode": "time", "t_final": 240e-6,
    }
    x = (np.arange(n) + 0.5) * (1.0 / n)
    st = np.zeros((9, n))
    hp = x <= 0.7
    st[0] = np.where(hp, ALPHA_FLOOR, 1 - ALPHA_FLOOR)
    st[1] = 50.0
    st[2] = np.where(hp, 1e9, 1e5)
    st[3] = 1 - st[0]
    st[4] = 1000.0
    st[5] = st[2]
    return case, st


# ---------------------------------------------------------------------------
# Batched and served workloads: independent two-phase Riemann problems
# ---------------------------------------------------------------------------
P_RIGHT = 1e5
RHO_AIR_TUBE = 50.0   # air density of the water/air shock tube (as shock_tube_1d)


def frozen_sound_speed(state_vals, eos=(AIR, WATER)):
    """Frozen mixture sound speed sqrt(sum_k Y_k c_k^2) of one cell (MixKapila.cpp:166-169)."""
    a = [state_vals[3 * k] for k in range(2)]
    r = [state_vals[3 * k + 1] for k in range(2)]
    p = [state_vals[3 * k + 2] for k in range(2)]
    rho = sum(a[k] * r[k] for k in range(2))
    c2 = 0.0
    for k, e in enumerate(eos):
        ck2 = e["gamma"] * (p[k] + e.get("pinf", 0.0)) / r[k]
        c2 += a[k] * r[k] / rho * ck2
    return math.sqrt(c2)


def riemann_case(dim, n, ratio, pos, precision="fp32s"):
    """Water (high pressure p = ratio * 1e5) against air (1e5 Pa), both at rest.

    1-D: water for x <= pos on [0, 1].  2-D: a water column of radius pos/2
    centred in [0, 1]^2.  Kapila model, MUSCL van Leer, transmissive
    boundaries, alpha floor 1e-6, CFL 0.8.  Runs to t_final = 0.3 * L / a_max,
    a_max the largest initial wave speed (frozen sound speed, fluid at rest).
    Returns (case, state, t_final)."""
    cells = [n, 1, 1] if dim == 1 else [n, n, 1]
    lengths = [1.0, 0.1, 0.1] if dim == 1 else [1.0, 1.0, 1.0]
    case = {"model": "kapila", "alpha_null": False, "eos": [AIR, WATER], "cells": cells, "lengths": lengths,
            "order": 2, "limiters": ["vanleer"] * 4, "cfl": 0.8, "precision": precision}
    pl = ratio * P_RIGHT
    left = [ALPHA_FLOOR, RHO_AIR_TUBE, pl, 1 - ALPHA_FLOOR, 1000.0, pl, 0.0, 0.0, 0.0]
    right = [1 - ALPHA_FLOOR, RHO_AIR_TUBE, P_RIGHT, ALPHA_FLOOR, 1000.0, P_RIGHT, 0.0, 0.0, 0.0]
    xc, yc, _ = _centres(cells, lengths)
    if dim == 1:
        inside = xc <= pos
    else:
        r = 0.5 * pos
        inside = ((xc[None, :] - 0.5) ** 2 + (yc[:, None] - 0.5) ** 2 <= r * r).ravel()
    st = np.empty((9, inside.size))
    for f in range(9):
        st[f] = np.where(inside, left[f], right[f])
    a_max = max(frozen_sound_speed(left), frozen_sound_speed(right))
    return case, st, 0.3 * lengths[0] / a_max


def batch_specs(count=2048, seed=0xEC0C0004, frac_1d=0.7, n1=(200, 50000), s2=(64, 768), ratio=(10.0, 1e4),
                pos=(0.2, 0.8)):
    """Parameters of the batched configuration: (dim, n, ratio, pos) per case,
    drawn with splitmix64 (log-uniform sizes and pressure ratios)."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        dim = 1 if rng.uniform() < frac_1d else 2
        lo, hi = n1 if dim == 1 else s2
        n = int(round(math.exp(math.log(lo) + rng.uniform() * (math.log(hi) - math.log(lo)))))
        r = math.exp(math.log(ratio[0]) + rng.uniform() * (math.log(ratio[1]) - math.log(ratio[0])))
        x = pos[0] + rng.uniform() * (pos[1] - pos[0])
        out.append((dim, n, r, x))
    return out


def bubble_request(seed, n=128, nb=8, rmin=4.0, rmax=16.0, precision="fp32s"):
    """3-D MultiP bubble case of the served stream (physics of bubbles_3d):
    ``nb`` air bubbles in water at 5e7 Pa, MUSCL minmod, stiff relaxation.
    Runs to t_final = 0.3 * L / a_max.  Returns (case, state, t_final)."""
    case, st = bubbles_3d(n=n, nb=nb, seed=seed, rmin=rmin, rmax=rmax, precision=precision)
    case.pop("time_mode", None)
    case.pop("steps", None)
    water = [ALPHA_FLOOR, 1.2, 5e7, 1 - ALPHA_FLOOR, 1000.0, 5e7]
    return case, st, 0.3 * case["lengths"][0] / frozen_sound_speed(water)


def request_stream(lam, count, seed=0xEC0C0005, frac_3d=0.05):
    """Poisson arrivals at rate ``lam`` per second: list of (arrival_s, kind, params).

    The sequence of requests depends on ``seed`` only; ``lam`` scales the
    exponential inter-arrival gaps, so streams at different rates carry the
    same requests.  kind "riemann": params (dim, n, ratio, pos) from the
    batched distribution; kind "bubbles": params = seed of a 3-D bubble case."""
    rng = SplitMix64(seed)
    gaps = SplitMix64(seed ^ 0x5EED0A11)
    t = 0.0
    out = []
    for i in range(count):
        t += -math.log(1.0 - gaps.uniform()) / lam
        if rng.uniform() < frac_3d:
            out.append((t, "bubbles", (seed + 7919 * (i + 1)) & 0xFFFFFFFF))
        else:
            out.append((t, "riemann", batch_specs(1, seed=rng.next())[0]))
    return out


def request_case(kind, params, precision="fp32s"):
    """(case, state, t_final) of one request of request_stream."""
    if kind == "riemann":
        return riemann_case(*params, precision=precision)
    return bubble_request(params, precision=precision)
