"""Batched and served solves (include/ecogen_pool.h) against solves run alone.

Every case admitted to a pool must give the same result as the same case run
alone through eco_create / eco_set_state / eco_run with the same precision:
bit for bit (the same device arithmetic on the same inputs), same step count
and physical time -- whether the pool replays hipGraphs or launches
directly, and whatever the mix and the admission slot of the other cases.
"""
import numpy as np
import pytest

from ecogen_amd import Context
from ecogen_amd import workloads as W

pytestmark = pytest.mark.gpu

# small mixed set: 1-D and 2-D Kapila water/air Riemann problems
SPECS = [(1, 150, 1e4, 0.7), (1, 700, 30.0, 0.25), (1, 2000, 800.0, 0.5), (2, 24, 1e3, 0.6), (2, 40, 50.0, 0.3),
         (2, 64, 5e3, 0.8), (1, 333, 10.0, 0.45), (2, 33, 200.0, 0.5)]


def _alone(case, st, t_final=0.0, nsteps=0):
    with Context(case) as ctx:
        ctx.set_state(st)
        s = ctx.run(nsteps=nsteps, t_final=t_final)
        return ctx.get_state(), s.steps, s.t


@pytest.fixture(scope="module")
def cases32():
    out = []
    for dim, n, r, x in SPECS:
        out.append(W.riemann_case(dim, n, r, x, precision="fp32s"))
    return out


@pytest.fixture(scope="module")
def alone32(cases32):
    return [_alone(c, st, t_final=tf) for c, st, tf in cases32]


def _pool_run(cases, **kw):
    from ecogen_amd.pool import Pool
    with Pool(hbm_budget=1 << 30, **kw) as pool:
        ids = [pool.admit(c, st, t_final=tf) for c, st, tf in cases]
        assert all(i is not None for i in ids)
        pool.run_all()
        res = [pool.result(i) for i in ids]
        for i in ids:
            pool.release(i)
        return res


def _same(res, alone):
    for (s, st), (a, steps, t) in zip(res, alone):
        assert st.steps == steps and st.t == t and st.done == 1
        assert np.array_equal(s, a), float(np.abs(s - a).max())


def test_batch_fp32s_matches_alone_graph(cases32, alone32):
    _same(_pool_run(cases32, use_graph=True, steps_per_chunk=24), alone32)


def test_batch_fp32s_matches_alone_direct(cases32, alone32):
    _same(_pool_run(cases32, use_graph=False, steps_per_chunk=12), alone32)


def test_batch_small_grid_matches(cases32, alone32):
    # a persistent grid far smaller than the tile count: every workgroup walks many tiles
    _same(_pool_run(cases32, use_graph=True, grid_blocks=7), alone32)


def test_batch_fp64_matches_alone():
    cases = [W.riemann_case(d, n, r, x, precision="fp64") for d, n, r, x in SPECS[:5]]
    alone = [_alone(c, st, t_final=tf) for c, st, tf in cases]
    _same(_pool_run(cases, use_graph=True), alone)


def test_iteration_mode_case():
    case, st, _ = W.riemann_case(2, 48, 300.0, 0.5, precision="fp32s")
    a, steps, t = _alone(case, st, nsteps=37)
    from ecogen_amd.pool import Pool
    with Pool(hbm_budget=1 << 28) as pool:
        cid = pool.admit(case, st, nsteps=37)
        pool.run_all()
        s, stt = pool.result(cid)
    assert stt.steps == 37 == steps and stt.t == t
    assert np.array_equal(s, a)


def test_staggered_admission(cases32, alone32):
    """Cases admitted while others run (different start slots) give the same results."""
    from ecogen_amd.pool import Pool
    with Pool(hbm_budget=1 << 30, steps_per_chunk=12) as pool:
        ids = {}
        for k, (c, st, tf) in enumerate(cases32):
            ids[k] = pool.admit(c, st, t_final=tf)
            pool.advance()  # one chunk between admissions
            pool.poll(keep=1)
        pool.run_all()
        res = [pool.result(ids[k]) for k in range(len(cases32))]
    _same(res, alone32)


def test_budget_admission():
    from ecogen_amd.pool import Pool
    case, st, tf = W.riemann_case(2, 64, 100.0, 0.5, precision="fp32s")
    need = Pool.case_bytes(case)
    # room for two cases (plus the FP64 admission staging), not three
    with Pool(hbm_budget=3 * need) as pool:
        a = pool.admit(case, st, t_final=tf)
        b = pool.admit(case, st, t_final=tf)
        assert a is not None and b is not None
        assert pool.admit(case, st, t_final=tf) is None
        pool.run_all()
        sa, _ = pool.result(a)
        sb, _ = pool.result(b)
        assert np.array_equal(sa, sb)
        pool.release(a)
        pool.poll(keep=0)
        c = pool.admit(case, st, t_final=tf)
        assert c is not None
        pool.run_all()
        sc, _ = pool.result(c)
        assert np.array_equal(sc, sa)


def test_serve_same_as_alone(cases32, alone32):
    """Requests arriving while others run: each served result equals the alone result."""
    from ecogen_amd.pool import Pool
    reqs = [{"case": c, "state": st, "t_final": tf, "arrival_s": 0.004 * k} for k, (c, st, tf) in
            enumerate(cases32)]
    with Pool(hbm_budget=1 << 30, steps_per_chunk=12) as pool:
        res, outs = pool.serve(reqs, keep_states=True)
    assert np.all(res[:, 1] >= res[:, 0])
    for k, (a, steps, t) in enumerate(alone32):
        assert int(res[k, 2]) == steps and res[k, 3] == t
        assert np.array_equal(outs[k], a)


def test_fp32s_batch_accuracy_vs_fp64(cases32, alone32):
    """FP32 storage against the FP64 path on the same cases: rel L1 <= 1e-5
    in mixture density, mixture pressure, velocity and volume fraction.

    Cases whose FP64 run hits the relaxation's 100-iteration limit
    (ECO_WARN_RELAX_NOT_CONVERGED: the air phase expanded to the IG pressure
    floor) are excluded: those cells hold non-converged states in either
    precision and do not measure the storage error."""
    checked = 0
    for (c, st, tf), (a32, steps32, _) in zip(cases32, alone32):
        c64 = dict(c, precision="fp64")
        with Context(c64) as ctx:
            ctx.set_state(st)
            s64 = ctx.run(nsteps=steps32)
            a64 = ctx.get_state()
        if s64.flags & 1:
            continue
        checked += 1
        nph = 2
        rho = lambda s: sum(s[3 * k] * s[3 * k + 1] for k in range(nph))
        p = lambda s: sum(s[3 * k] * s[3 * k + 2] for k in range(nph))
        for name, f in (("rho", rho), ("p", p), ("alpha", lambda s: s[0]), ("u", lambda s: s[6])):
            x, y = f(a32), f(a64)
            den = np.abs(y).sum()
            err = np.abs(x - y).sum() / den if den > 0 else 0.0
            assert err <= 1e-5, (c["cells"], name, err)
    assert checked >= 4


def test_batch_fp64_boundaries_match_alone():
    """FP64 decks with walls, outflow, injection and tank boundaries (first
    order and MUSCL, 1-D and 2-D), batched with a transmissive Riemann case:
    every result bit-identical to the same case run alone."""
    from conftest import initial_state, load_golden
    from ecogen_amd.pool import Pool
    names = ["c1_kapila_fo_wall", "kapila_1d_o2_outflow", "kapila_1d_o2_injection", "multip_1d_o2_relax_tank",
             "euler_1d_o2_injection", "kapila_2d_fo_tank", "kapila_2d_o2_wall"]
    cases = []
    for nm in names:
        case, d = load_golden(nm)
        st = initial_state(case, d)
        n = case["steps"] if case["time_mode"] == "iterations" else 0
        cases.append((dict(case, precision="fp64"), st, n, case.get("t_final", 0.0)))
    rc, rst, _ = W.riemann_case(2, 64, 100.0, 0.4, precision="fp64")
    cases.append((rc, rst, 40, 0.0))
    print("pool of", len(cases), "FP64 cases with boundaries", flush=True)
    alone = [_alone(c, st, nsteps=n, t_final=tf) for c, st, n, tf in cases]
    with Pool(hbm_budget=1 << 30, steps_per_chunk=12) as pool:
        ids = [pool.admit(c, st, nsteps=n, t_final=tf) if n else pool.admit(c, st, t_final=tf) for c, st, n, tf in cases]
        assert all(i is not None for i in ids)
        pool.run_all()
        res = [pool.result(i) for i in ids]
    _same(res, alone)


def test_batch_three_and_four_phases_match_alone():
    """Three- and four-phase FP64 decks (their own kernel classes) batched
    with a two-phase case: every result bit-identical to its single solve."""
    from conftest import initial_state, load_golden
    from ecogen_amd.pool import Pool
    cases = []
    for nm in ["kapila_1d_o2_3ph", "multip_3d_o2_relax_3ph", "kapila_1d_o2_4ph", "multip_2d_o2_relax_4ph",
               "kapila_1d_o2_vanleer"]:
        case, d = load_golden(nm)
        cases.append((dict(case, precision="fp64"), initial_state(case, d), case["steps"]))
    print("pool of 2-, 3- and 4-phase FP64 cases", flush=True)
    alone = [_alone(c, st, nsteps=n) for c, st, n in cases]
    with Pool(hbm_budget=1 << 30, steps_per_chunk=12) as pool:
        ids = [pool.admit(c, st, nsteps=n) for c, st, n in cases]
        assert all(i is not None for i in ids)
        pool.run_all()
        res = [pool.result(i) for i in ids]
    _same(res, alone)
