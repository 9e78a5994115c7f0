"""Resumed runs (<resumeSimulation fileNumber="N"/>, Run::resumeSimulation,
src/Run.cpp:164-190) against the reference's own resumed runs.

Each fixture under tests/golden/resume holds what the reference's first run
wrote (result file N and infoCalcul.out) and the final state and
infoCalcul.out rows of its second, resumed run.  The run directory is rebuilt
from them, the deck reader takes the state, iteration, time and dt back from
the files, and the march continues:

* the oracle (CPU) from that point must reproduce the reference's resumed
  run bit for bit: step count, final time and every printed field.  The
  resumed run's first step uses the dt read from infoCalcul.out, so it ends
  elsewhere than the uninterrupted run (by up to 5e-5 here); the test checks
  that the fixture shows that difference, so it pins the restart clock;
* the device (GPU, through python -m ecogen_amd.deck --vtk) must match it to
  the parity bar and continue infoCalcul.out and the file numbering as the
  reference does.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT, compared_fields, parity_quantities, rel_l1
from deck import write_deck

import oracle as O

RES = ROOT / "tests" / "golden" / "resume"
NAMES = sorted(p.stem for p in RES.glob("*.npz"))


def load_resume(name):
    d = np.load(RES / f"{name}.npz")
    return json.loads(bytes(d["case_json"]).decode()), d


def rebuild_rundir(name, case, d, root: Path):
    """The reference's run directory after its first run: the deck with
    resumeSimulation, result file N and infoCalcul.out."""
    write_deck(dict(case, name=name), root)
    rdir = root / "results" / name
    rdir.mkdir(parents=True)
    (rdir / f"result_CPU0_AMR0_TIME{case['resume']}.vtr").write_bytes(bytes(d["resume_vtr"]))
    (rdir / "infoCalcul.out").write_bytes(bytes(d["info_run1"]))
    return rdir


def _final_errors(case, d, out, prefix="final"):
    fin = O.state_from_golden(case, d, prefix)
    errs = {}
    for f, nm in compared_fields(case):
        if case["model"] == "kapila" and nm.startswith("p"):
            continue  # Kapila prints the mixture pressure only
        errs[nm] = rel_l1(out[f], fin[f])
    if case["model"] == "kapila":
        nph = len(case["eos"])
        errs["p_mix"] = rel_l1(sum(out[3 * k] * out[3 * k + 2] for k in range(nph)), d[f"{prefix}_p_mix"])
    return errs


@pytest.mark.parametrize("name", NAMES)
def test_oracle_resumes_like_reference(name, tmp_path):
    from ecogen_amd import deck
    case, d = load_resume(name)
    rebuild_rundir(name, case, d, tmp_path / "run")
    got, st0 = deck.load(tmp_path / "run")
    assert got["resume"] == case["resume"]
    state, it0, t0, dt0, kept = deck.resume_point(got, tmp_path / "run")
    info = d["info"]
    assert it0 == int(info[case["resume"], 1]) and t0 == info[case["resume"], 2] and dt0 == info[case["resume"], 4]
    assert len(kept) == case["resume"] + 2  # the CPU count, rows 0..N
    if case["time_mode"] == "iterations":
        out, it, t, _, _ = O.run(case, state, nsteps=case["steps"] - it0, resume=(t0, dt0))
    else:
        out, it, t, _, _ = O.run(case, state, t_final=case["t_final"], resume=(t0, dt0))
    assert it0 + it == int(info[-1, 1])
    assert t == info[-1, 2]
    errs = _final_errors(case, d, out)
    assert max(errs.values()) == 0.0, errs
    # the restart clock matters: the uninterrupted run ends elsewhere
    assert max(_final_errors(case, d, out, "run1_final").values()) > 1e-13


def test_result_reader_binary_and_ascii(tmp_path):
    """read_vtr reads back what the writer wrote: ASCII at 17 digits exactly,
    binary as Float32."""
    from cases import CASES
    from ecogen_amd import deck, output
    case = CASES["multip_1d_o2_relax"]
    rng = np.random.default_rng(3)
    n = case["cells"][0]
    st = np.zeros((9, n))
    st[0] = rng.uniform(0.1, 0.9, n); st[3] = 1 - st[0]
    st[1], st[4] = rng.uniform(1, 60, n), rng.uniform(900, 1100, n)
    st[2], st[5] = rng.uniform(1e5, 1e9, n), rng.uniform(1e5, 1e9, n)
    st[6] = rng.uniform(-50, 50, n)
    for binary in (False, True):
        f = tmp_path / f"r{int(binary)}.vtr"
        output.write_vtr(f, case, st, binary=binary, precision=17)
        arrs = deck.read_vtr(f)
        for k in range(2):
            a = arrs[[x for x in arrs if x.startswith(f"F{k}_Pressure_")][0]]
            want = st[3 * k + 2].astype(np.float32).astype(np.float64) if binary else st[3 * k + 2]
            assert np.array_equal(a, want)
        assert arrs["Velocity_Mixture"].shape == (n, 3)


def test_resume_refusals(tmp_path):
    from ecogen_amd import deck
    name = NAMES[0]
    case, d = load_resume(name)
    rdir = rebuild_rundir(name, case, d, tmp_path / "a")
    got, _ = deck.load(tmp_path / "a")
    with pytest.raises(deck.DeckError, match="check file"):
        deck.resume_point(dict(got, resume=case["resume"] + 50), tmp_path / "a")
    # as the reference: no resume from binary or GNU results (Output.cpp:279, Output.h:79)
    with pytest.raises(deck.DeckError, match="binary"):
        deck.resume_point(dict(got, output=dict(got["output"], binary=True)), tmp_path / "a")
    with pytest.raises(deck.DeckError, match="output format"):
        deck.resume_point(dict(got, output=dict(got["output"], format="GNU")), tmp_path / "a")
    text = (rdir / "infoCalcul.out").read_text().split("\n")
    (rdir / "infoCalcul.out").write_text("\n".join(["2"] + text[1:]))
    with pytest.raises(deck.DeckError, match="number of CPU"):
        deck.resume_point(got, tmp_path / "a")


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_deck_resumes_like_reference(name, tmp_path):
    """python -m ecogen_amd.deck --vtk on the reference's run directory after
    its first run: the final state within the parity bar (1e-12 relative L1),
    and infoCalcul.out truncated after row N and continued as the reference
    continued it (iterations exact, time and dt to 1e-12)."""
    from ecogen_amd import deck
    case, d = load_resume(name)
    print(f"case {name}: resume from file {case['resume']}", flush=True)
    rdir = rebuild_rundir(name, case, d, tmp_path / "run")
    deck.main([str(tmp_path / "run"), "--out", str(tmp_path / "final.npz"), "--vtk"])
    r = np.load(tmp_path / "final.npz")
    info = d["info"]
    assert int(r["steps"]) == int(info[-1, 1])
    fin = O.state_from_golden(case, d, "final")
    got, want = parity_quantities(case, r["state"]), parity_quantities(case, fin)
    for nm in got:
        if nm == "p" and case["model"] == "kapila":
            e = rel_l1(got[nm], d["final_p_mix"])
        elif nm == "p" and case["model"] == "multip" and case.get("relax_p"):
            continue  # the reference prints p* of the relaxation, not sum alpha_k p_k
        else:
            e = rel_l1(got[nm], want[nm])
        assert e <= 1e-12, f"{nm}: {e:.3e}"
    rows = [ln.split() for ln in (rdir / "infoCalcul.out").read_text().split("\n")[1:] if ln.strip()]
    assert len(rows) == len(info)
    for row, ref in zip(rows, info):
        assert int(row[0]) == int(ref[0]) and int(row[1]) == int(ref[1])
        assert float(row[2]) == pytest.approx(ref[2], rel=1e-12, abs=0)
        assert float(row[4]) == pytest.approx(ref[4], rel=1e-12, abs=0)
    last = int(info[-1, 0])
    assert (rdir / f"result_CPU0_AMR0_TIME{last}.vtr").exists()
