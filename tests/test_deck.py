"""The deck reader (ecogen_amd/deck.py) against the reference's own inputs:
every golden case's deck, written in the reference schema exactly as it was
given to the reference executable, reads back to the same configuration and
to the t = 0 state the reference printed for it (bit for bit)."""
import numpy as np
import pytest

from conftest import golden_names, load_golden

import oracle as O
from deck import write_deck



@pytest.mark.parametrize("name", golden_names())
def test_deck_reads_back_the_reference_inputs(name, tmp_path):
    from ecogen_amd import deck
    case, d = load_golden(name)
    write_deck(dict(case, name=name), tmp_path / "run")
    got, state = deck.load(tmp_path / "run")
    for key in ("model", "order", "limiters", "cells", "lengths", "cfl", "time_mode"):
        assert got[key] == case[key], key
    assert got.get("steps") == case.get("steps") if case["time_mode"] == "iterations" else got["t_final"] == case["t_final"]
    assert got["alpha_null"] == bool(case.get("alpha_null", False))
    assert got["relax_p"] == bool(case.get("relax_p", False) and case["model"] == "multip")
    assert got.get("relax_pt", False) == bool(case.get("relax_pt", False))
    dim = sum(1 for c in case["cells"] if c > 1)
    assert got["bc"][:2 * dim] == case.get("bc", ["abs"] * 6)[:2 * dim]
    for i, kind in enumerate(got["bc"][:2 * dim]):
        if kind == "outflow":
            assert got["bc_p0"][i] == case["bc_p0"][i]
        if kind == "injection":
            assert got["bc_inj"][i] == case["bc_inj"][i]
        if kind == "tank":
            assert got["bc_tank"][i] == case["bc_tank"][i]
        if kind in ("outflow", "injection", "tank") and case.get("transports"):
            assert got["bc_tr"][i] == case["bc_tr"][i]
    assert got.get("transports", []) == case.get("transports", [])
    for a, b in zip(got["eos"], case["eos"]):
        for k in ("type", "gamma", "cv"):
            assert a[k] == b[k]
        assert a["pinf"] == b.get("pinf", 0.0) and a["eref"] == b.get("eref", 0.0)
    want = O.state_from_golden(case, d)
    assert state.shape[1] == want.shape[1]
    n = min(state.shape[0], want.shape[0])
    if case["model"] == "kapila":
        # the reference prints the mixture pressure sum alpha_k p_k, not p_k
        nph = len(case["eos"])
        pk = [3 * k + 2 for k in range(nph)]
        keep = [f for f in range(n) if f not in pk]
        assert np.array_equal(state[keep], want[keep])
        pm = sum(state[3 * k] * state[3 * k + 2] for k in range(nph))
        assert np.allclose(pm, d["init_p_mix"], rtol=1e-15, atol=0)
    else:
        assert np.array_equal(state[:n], want[:n])


def test_deck_refusals(tmp_path):
    from ecogen_amd import deck
    case, _ = load_golden("kapila_1d_o2_vanleer")
    # every type Input.cpp:849-856 reads is read; others are refused as there
    write_deck(dict(case, name="x", bc=["periodic"] * 6), tmp_path / "a")
    with pytest.raises(deck.DeckError, match="periodic"):
        deck.load(tmp_path / "a")
    write_deck(dict(case, name="x", bc=["symmetry"] * 6), tmp_path / "b")
    assert deck.load(tmp_path / "b")[0]["bc"][:2] == ["wall"] * 2  # a 1-D deck names X- and X+ only
    inj = {"m0": 10.0, "alpha": [0.5, 0.4], "rho": [1.0, 1000.0], "p": [1e5, 1e5]}
    write_deck(dict(case, name="x", bc=["injection"] * 6, bc_inj=[inj] * 6), tmp_path / "c")
    with pytest.raises(deck.DeckError, match="sum of alpha"):
        deck.load(tmp_path / "c")
    tank = {"p0": 2e5, "T0": 300.0, "massFrac": [0.5, 0.6]}
    write_deck(dict(case, name="x", bc=["tank"] * 6, bc_tank=[tank] * 6), tmp_path / "d")
    with pytest.raises(deck.DeckError, match="sum of massFrac"):
        deck.load(tmp_path / "d")
    write_deck(dict(case, name="x", output='format="GNU" binary="true"'), tmp_path / "e")
    with pytest.raises(deck.DeckError, match="binary"):
        deck.load(tmp_path / "e")
    write_deck(dict(case, name="x"), tmp_path / "f")
    mx = tmp_path / "f" / "case" / "mainV5.xml"
    mx.write_text(mx.read_text().replace("</computationParam>", '<cut1D><vertex x="0" y="0" z="0"/>'
                                         '<vecDir x="0" y="0" z="0"/></cut1D></computationParam>'))
    with pytest.raises(deck.DeckError, match="null vecDir"):  # GOLine::GOLine
        deck.load(tmp_path / "f")
    (tmp_path / "a" / "case" / "meshV5.xml").write_text(
        '<?xml version="1.0"?><mesh><type structure="unstructured"/></mesh>')
    with pytest.raises(deck.DeckError, match="cartesian"):
        deck.load(tmp_path / "a")
    # transports: the Euler model has none (Model.h:127); numberTransports
    # needs as many named <transport> elements (Input.cpp:528-540); an outflow
    # side needs a value for each (BoundCondOutflow.cpp:79)
    eu, _ = load_golden("euler_1d_o2")
    write_deck(dict(eu, name="x", transports=["c"], states=[dict(s, tr=[1.0]) for s in eu["states"]]), tmp_path / "g")
    with pytest.raises(deck.DeckError, match="Euler"):
        deck.load(tmp_path / "g")
    tr, _ = load_golden("kapila_1d_fo_transport_outflow")
    write_deck(dict(tr, name="x"), tmp_path / "h")
    mx = tmp_path / "h" / "case" / "modelV4.xml"
    mx.write_text(mx.read_text().replace('numberTransports="1"', 'numberTransports="2"'))
    with pytest.raises(deck.DeckError, match="numberTransports"):
        deck.load(tmp_path / "h")
    write_deck(dict(tr, name="x", bc_tr=[[4.0], None] + [None] * 4), tmp_path / "i")
    with pytest.raises(deck.DeckError, match="not enough transport"):
        deck.load(tmp_path / "i")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c1_kapila_fo_wall", "kapila_2d_fo_wall", "multip_1d_o2_relax",
                                  "kapila_1d_o2_injection", "multip_1d_o2_relax_tank",
                                  "kapila_1d_o2_transport_wall_tank", "multip_2d_fo_relax_transport",
                                  "thermaleq_1d_fo_outflow", "thermaleq_2d_o2", "eulerhom_2d_fo",
                                  "kapila_1d_o2_ptmu"])
def test_deck_runs_on_the_device(name, tmp_path):
    """python -m ecogen_amd.deck on a reference deck: the final state matches
    the reference's output (1e-12 relative L1, the parity bar; 1e-11 with the
    liquid-vapour saturation laws, whose C library log misrounds where the
    device's is correctly rounded)."""
    from conftest import parity_quantities, rel_l1
    from ecogen_amd import deck
    case, d = load_golden(name)
    print(f"case {name} from its deck", flush=True)
    write_deck(dict(case, name=name), tmp_path / "run")
    out = tmp_path / "final.npz"
    deck.main([str(tmp_path / "run"), "--out", str(out)])
    r = np.load(out)
    fin = O.state_from_golden(case, d, "final")
    got, want = parity_quantities(case, r["state"]), parity_quantities(case, fin)
    for nm in got:
        if nm == "p" and case["model"] == "kapila":
            e = rel_l1(got[nm], d["final_p_mix"])
        elif nm == "p" and case["model"] == "multip" and case.get("relax_p"):
            continue  # the reference prints p* of the relaxation, not sum alpha_k p_k
        else:
            e = rel_l1(got[nm], want[nm])
        sat = case["model"] == "eulerhomogeneous" or case.get("relax_ptmu")
        assert e <= (1e-11 if sat else 1e-12), f"{nm}: {e:.3e}"


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["kapila_2d_fo_wall_6x4", "euler_1d_fo_30", "kapila_1d_o2_transport_30",
                                  "thermaleq_1d_o2_30"])
def test_deck_vtk_results_match_reference_files(name, tmp_path):
    """python -m ecogen_amd.deck --vtk writes the reference's result files:
    the collection file character for character, and the last result file
    with the reference's arrays and header, values within 1e-10 relative of
    the reference's own file (device and reference differ in last bits)."""
    import json
    from pathlib import Path
    from ecogen_amd import deck
    from test_output import parse
    d = Path(__file__).resolve().parent / "golden" / "vtk" / name
    case = json.loads((d / "case.json").read_text())
    print(f"case {name} deck --vtk", flush=True)
    write_deck(dict(case, name=name, output='format="XML" binary="false" precision="17"'), tmp_path / "run")
    deck.main([str(tmp_path / "run"), "--out", str(tmp_path / "f.npz"), "--vtk"])
    res = tmp_path / "run" / "results" / name
    assert (res / "collection.pvd").read_text() == (d / "collection.pvd").read_text()
    ours_text = (res / "result_CPU0_AMR0_TIME1.vtr").read_text()
    ref_text = (d / "final.vtr").read_text()
    assert ours_text[: ours_text.index("<CellData>")] == ref_text[: ref_text.index("<CellData>")]
    ours, ref = parse(ours_text), parse(ref_text)
    assert [(n, c) for n, c, _ in ours] == [(n, c) for n, c, _ in ref]
    for (n, _, a), (_, _, b) in zip(ours, ref):
        scale = max(np.abs(b).max(), 1e-300)
        assert np.abs(a - b).max() <= 1e-10 * scale, n


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c1_kapila_fo_wall", "kapila_1d_o2_wall"])
def test_deck_info_file_matches_reference(name, tmp_path):
    """infoCalcul.out of a --vtk run against the reference's own (kept in the
    golden): result number, iteration, physical time and last dt per output
    (the computation-time column is a wall-clock time and is not compared)."""
    from ecogen_amd import deck
    case, d = load_golden(name)
    print(f"case {name} deck --vtk info", flush=True)
    write_deck(dict(case, name=name), tmp_path / "run")
    deck.main([str(tmp_path / "run"), "--out", str(tmp_path / "f.npz"), "--vtk"])
    lines = (tmp_path / "run" / "results" / name / "infoCalcul.out").read_text().split("\n")
    assert lines[0] == "1"
    ours = np.array([[float(x) for x in r.split()] for r in lines[1:] if r.strip()])
    ref = d["info"]
    assert ours.shape == ref.shape
    assert np.array_equal(ours[:, :2], ref[:, :2])
    for col in (2, 4):
        assert np.allclose(ours[:, col], ref[:, col], rtol=1e-12, atol=0.0), col


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["kapila_2d_fo_wall_6x4", "euler_1d_fo_30"])
def test_deck_gnu_results_match_reference_files(name, tmp_path):
    """python -m ecogen_amd.deck --vtk on the deck with GNU output writes the
    reference's result file and visualisation.gnu: the script character for
    character, the rows with the reference's columns and every value equal
    to its 6 printed digits (5e-6 relative; device and reference differ in
    last bits, which 6 digits hide but for a rounding tie)."""
    import json
    from pathlib import Path
    from ecogen_amd import deck
    d = Path(__file__).resolve().parent / "golden" / "vtk" / name
    case = json.loads((d / "case.json").read_text())
    print(f"case {name} deck --vtk with GNU output", flush=True)
    write_deck(dict(case, name=name, output='format="GNU" binary="false"'), tmp_path / "run")
    deck.main([str(tmp_path / "run"), "--out", str(tmp_path / "f.npz"), "--vtk"])
    res = tmp_path / "run" / "results" / name
    assert (res / "visualisation.gnu").read_text() == (d / "visualisation.gnu").read_text()
    ours = [ln.split() for ln in (res / "result_CPU0_TIME1.out").read_text().split("\n") if ln.strip()]
    ref = [ln.split() for ln in (d / "final.out").read_text().split("\n") if ln.strip()]
    assert [len(r) for r in ours] == [len(r) for r in ref]
    a, b = np.array(ours, dtype=float), np.array(ref, dtype=float)
    assert np.all(np.abs(a - b) <= 5e-6 * np.abs(b) + 1e-300)
    assert (res / "infoCalcul.out").exists() and not (res / "collection.pvd").exists()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["kapila_2d_fo_probes", "euler_1d_fo_probes"])
def test_deck_probe_files_match_reference(name, tmp_path):
    """python -m ecogen_amd.deck --vtk on a deck with probes: each probe file
    has the reference's rows (times and values equal to their 6 printed
    digits; device and reference differ in last bits) and its script."""
    import json
    from pathlib import Path
    from ecogen_amd import deck
    d = Path(__file__).resolve().parent / "golden" / "probes" / name
    case = json.loads((d / "case.json").read_text())
    print(f"case {name} deck --vtk with probes", flush=True)
    write_deck(dict(case, name=name), tmp_path / "run")
    deck.main([str(tmp_path / "run"), "--out", str(tmp_path / "f.npz"), "--vtk"])
    res = tmp_path / "run" / "results" / name / "probes"
    for pr in case["probes"]:
        assert (res / f"visualisation{pr['name']}.gnu").read_text() == (d / f"visualisation{pr['name']}.gnu").read_text()
        ours = [ln.split() for ln in (res / f"{pr['name']}.out").read_text().split("\n") if ln.strip()]
        ref = [ln.split() for ln in (d / f"{pr['name']}.out").read_text().split("\n") if ln.strip()]
        assert [len(r) for r in ours] == [len(r) for r in ref]
        a, b = np.array(ours, dtype=float), np.array(ref, dtype=float)
        assert np.all(np.abs(a - b) <= 5e-6 * np.abs(b) + 1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["kapila_2d_fo_cuts", "euler_1d_fo_cuts", "kapila_3d_o2_cuts"])
def test_deck_cut_files_match_reference(name, tmp_path):
    """python -m ecogen_amd.deck --vtk on a deck with cuts writes the
    reference's cut files (the same cells and coordinates, values equal to
    their 6 printed digits) and scripts under results/<run>/cuts."""
    import json
    from pathlib import Path
    from ecogen_amd import deck
    d = Path(__file__).resolve().parent / "golden" / "cuts" / name
    case = json.loads((d / "case.json").read_text())
    print(f"case {name} deck --vtk with cuts", flush=True)
    write_deck(dict(case, name=name), tmp_path / "run")
    deck.main([str(tmp_path / "run"), "--out", str(tmp_path / "f.npz"), "--vtk"])
    res = tmp_path / "run" / "results" / name / "cuts"
    assert sorted(f.name for f in res.iterdir()) == sorted(f.name for f in d.iterdir() if f.name != "case.json")
    for f in sorted(d.glob("*.gnu")):
        assert (res / f.name).read_text() == f.read_text()
    for f in sorted(d.glob("*.out")):
        ours = [ln.split() for ln in (res / f.name).read_text().split("\n") if ln.strip()]
        ref = [ln.split() for ln in f.read_text().split("\n") if ln.strip()]
        assert [len(r) for r in ours] == [len(r) for r in ref], f.name
        a, b = np.array(ours, dtype=float), np.array(ref, dtype=float)
        assert np.all(np.abs(a - b) <= 5e-6 * np.abs(b) + 1e-300), f.name


@pytest.mark.gpu
def test_deck_iteration_output_only_at_multiples_of_iterfreq(tmp_path):
    """Iteration control writes a result file at t = 0 and whenever the
    iteration is a multiple of iterFreq (Run.cpp:239-241), and none at a last
    iteration that is not: 25 steps at iterFreq 10 give TIME0..TIME2 and three
    infoCalcul.out rows (iterations 0, 10, 20) after the CPU count, while the
    run still ends at iteration 25."""
    from ecogen_amd import deck
    case, _ = load_golden("c1_kapila_fo_wall")
    case = dict(case, name="freq", time_mode="iterations", steps=25, iter_freq=10)
    case.pop("t_final", None)
    print("case c1_kapila_fo_wall deck --vtk, 25 steps at iterFreq 10", flush=True)
    write_deck(case, tmp_path / "run")
    deck.main([str(tmp_path / "run"), "--out", str(tmp_path / "f.npz"), "--vtk"])
    res = tmp_path / "run" / "results" / "freq"
    assert sorted(p.name for p in res.glob("result_*.vtr")) == [f"result_CPU0_AMR0_TIME{k}.vtr" for k in range(3)]
    lines = [ln for ln in (res / "infoCalcul.out").read_text().split("\n") if ln.strip()]
    assert lines[0] == "1" and [int(r.split()[1]) for r in lines[1:]] == [0, 10, 20]
    assert int(np.load(tmp_path / "f.npz")["steps"]) == 25
