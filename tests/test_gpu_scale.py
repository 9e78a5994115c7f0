"""The batched and served configurations at their own scale.

Batched: all 2048 water/air Riemann problems of the batched workload (seed
0xEC0C0004: 70 % 1-D with 200 to 50,000 cells, 30 % 2-D with sides 64 to
768), each run to its own t_final in one pool, in the precision the routing
rule gives it.  Every 16th case is run alone: the pooled result must equal
it bit for bit, with the same step count and time.  Every case the rule
keeps in FP32 storage must lie within 1e-5 relative L1 of its FP64 result
(the FP64 ones are the FP64 result).

Served: one Poisson stream at 200 requests/s drawn from the same
distribution plus 5 % 3-D 128^3 bubble cases (MultiP with relaxation),
through the serving loop with a 200 GB HBM budget: every served result must
equal the same request solved alone, bit for bit.
"""
import numpy as np
import pytest

from ecogen_amd import Context
from ecogen_amd import workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


def _alone(case, st, t_final):
    with Context(case) as ctx:
        ctx.set_state(st)
        s = ctx.run(t_final=t_final)
        return ctx.get_state(), s


def _l1(case, a, b):
    nph = len(case["eos"])
    dim = sum(1 for n in case["cells"] if n > 1)
    qa = [sum(a[3 * k] * a[3 * k + 1] for k in range(nph)), sum(a[3 * k] * a[3 * k + 2] for k in range(nph)),
          np.concatenate([a[3 * nph + d] for d in range(dim)])] + [a[3 * k] for k in range(nph)]
    qb = [sum(b[3 * k] * b[3 * k + 1] for k in range(nph)), sum(b[3 * k] * b[3 * k + 2] for k in range(nph)),
          np.concatenate([b[3 * nph + d] for d in range(dim)])] + [b[3 * k] for k in range(nph)]
    return max(float(np.abs(x - y).sum() / np.abs(y).sum()) for x, y in zip(qa, qb))


def test_batched_2048_cases_routed_equal_alone():
    from ecogen_amd.pool import Pool, route_precision
    specs = W.batch_specs(2048)
    cases = []
    for sp in specs:
        c, st, tf = W.riemann_case(*sp)
        cases.append((dict(c, precision=route_precision(c, st, tf)), st, tf))
    n32 = sum(1 for c, _, _ in cases if c["precision"] == "fp32s")
    print(f"2048 cases: {n32} in FP32 storage, {len(cases) - n32} in FP64", flush=True)
    with Pool(hbm_budget=96 << 30, steps_per_chunk=24) as pool:
        ids = [pool.admit(c, st, t_final=tf) for c, st, tf in cases]
        assert all(i is not None for i in ids)
        pool.run_all()
        res = [pool.result(i) for i in ids]
    print("pool done", flush=True)
    for k in range(0, len(cases), 16):
        c, st, tf = cases[k]
        a, s = _alone(c, st, tf)
        p, ps = res[k]
        assert (ps.steps, ps.t, ps.flags) == (s.steps, s.t, s.flags), k
        assert np.array_equal(p, a), f"case {k} ({c['precision']}) differs from its solve alone"
    print("every 16th case equals its solve alone", flush=True)
    worst = 0.0
    for k, (c, st, tf) in enumerate(cases):
        if c["precision"] != "fp32s":
            continue
        a64, s64 = _alone(dict(c, precision="fp64"), st, tf)
        assert res[k][1].steps == s64.steps
        e = _l1(c, res[k][0], a64)
        worst = max(worst, e)
        assert e <= 1e-5, f"case {k}: FP32 storage {e:.3e} from FP64"
    print(f"FP32-storage cases within 1e-5 of FP64 (worst {worst:.2e})", flush=True)


def test_served_stream_with_3d_bubbles_equal_alone():
    from ecogen_amd.pool import Pool, route_precision
    base = W.request_stream(200.0, 120)
    assert any(kind == "bubbles" for _, kind, _ in base)
    reqs = []
    for arr, kind, prm in base:
        c, st, tf = W.request_case(kind, prm)
        reqs.append({"case": dict(c, precision=route_precision(c, st, tf)), "state": st, "t_final": tf,
                     "arrival_s": arr})
    print(f"120 requests, {sum(1 for _, k, _ in base if k == 'bubbles')} 3-D 128^3 bubble cases", flush=True)
    with Pool(hbm_budget=200 << 30, steps_per_chunk=24) as pool:
        res, outs = pool.serve(reqs, keep_states=True)
    print("served", flush=True)
    for i, r in enumerate(reqs):
        a, s = _alone(r["case"], r["state"], r["t_final"])
        assert int(res[i, 2]) == s.steps and res[i, 3] == s.t, i
        assert np.array_equal(outs[i], a), f"request {i} differs from its solve alone"
