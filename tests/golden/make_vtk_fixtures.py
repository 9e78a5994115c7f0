"""Reference VTK output files for the output-writer tests.

Runs the reference executable (``oracle/_ref/ECOGEN``, built by
``oracle/ref.mk``) on small decks and keeps its last result file (ASCII,
17 digits), its ``collection.pvd`` and the case description, under
``tests/golden/vtk/<case>/``; for 1-D and 2-D decks also the last result file
and ``visualisation.gnu`` of the same deck with GNU output.  The files are the
reference's outputs: data.

    python tests/golden/make_vtk_fixtures.py
"""
from __future__ import annotations

import json
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(REPO / "oracle"))
from cases import CASES, _bubbles_3d, _euler_1d, _shock_tube_1d, _tank, _thermaleq, _with_transports  # noqa: E402
from deck import write_deck  # noqa: E402

REF_BIN = REPO / "oracle" / "_ref" / "ECOGEN"

FIXTURES = {
    "kapila_2d_fo_wall_6x4": dict(CASES["kapila_2d_fo_wall"], cells=[6, 4, 1], steps=3),
    "kapila_1d_o2_alphanull_30": _shock_tube_1d("kapila", 2, ["vanleer"] * 4, 12, cells=30, alpha_null=True),
    "multip_1d_o2_relax_40": _shock_tube_1d("multip", 2, ["minmod"] * 4, 15, cells=40, relax=True),
    "euler_1d_fo_30": _euler_1d(1, ["none"] * 4, 20, cells=30),
    "kapila_3d_o2_6": _bubbles_3d(n=6, steps=3, nb=1, model="kapila", relax=False, limiters=("vanleer",) * 4),
    "kapila_1d_fo_outflow_30": dict(_shock_tube_1d("kapila", 1, ["none"] * 4, 20, cells=30),
                                    bc=["outflow", "outflow"] + ["abs"] * 4, bc_p0=[2e9, 5e4] + [0.0] * 4),
    # transported scalars: T1, T2 arrays (OutputXML.cpp:435-447), Transport<k> columns
    "kapila_1d_o2_transport_30": _with_transports(
        dict(_shock_tube_1d("kapila", 2, ["minmod", "vanleer", "minmod", "vanleer"], 15, cells=30),
             bc=["wall", "tank"] + ["abs"] * 4, bc_tank=_tank(2, 3e5, 300.0, 1)),
        ["color", "dye"], {"LP": [1.0, 0.5], "HP": [0.0, 2.0]}, bc_tr=[None, [5.0, -1.0]] + [None] * 4),
    # the thermal-equilibrium model's arrays (Alpha, Density; mixture density,
    # pressure, temperature, Velocity_Mixture_norm)
    "thermaleq_1d_o2_30": _thermaleq(_shock_tube_1d("kapila", 2, ["vanleer"] * 4, 12, cells=30)),
    "thermaleq_2d_fo_6x4": _thermaleq(dict(CASES["kapila_2d_fo_wall"], cells=[6, 4, 1], steps=3)),
    "multip_2d_fo_transport_6x4": _with_transports(
        dict(CASES["multip_2d_fo_relax"], cells=[6, 4, 1], steps=3), ["color"], {"water": [1.0], "shocked": [0.25]}),
}


def main(argv):
    for name, case in FIXTURES.items():
        if argv[1:] and name not in argv[1:]:
            continue
        tmp = Path(tempfile.mkdtemp(prefix="ecogen_vtk_"))
        try:
            write_deck(dict(case, name=name), tmp)
            r = subprocess.run([str(REF_BIN)], cwd=tmp, capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
            rdir = tmp / "results" / name
            vtrs = sorted(rdir.glob("result_CPU0_AMR0_TIME*.vtr"), key=lambda p: int(p.stem.rsplit("TIME", 1)[1]))
            out = HERE / "vtk" / name
            out.mkdir(parents=True, exist_ok=True)
            shutil.copy(vtrs[-1], out / "final.vtr")
            shutil.copy(rdir / "collection.pvd", out / "collection.pvd")
            (out / "case.json").write_text(json.dumps(case))
            print(f"{name}: {vtrs[-1].name} -> {out}")
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        if sum(1 for c in case["cells"] if c > 1) == 3:
            continue  # the reference writes no gnuplot script in 3-D (OutputGNU.cpp:110)
        # the same deck with GNU output (OutputGNU): the run is deterministic, so
        # its last file holds the state of final.vtr
        tmp = Path(tempfile.mkdtemp(prefix="ecogen_gnu_"))
        try:
            write_deck(dict(case, name=name, output='format="GNU" binary="false"'), tmp)
            r = subprocess.run([str(REF_BIN)], cwd=tmp, capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
            rdir = tmp / "results" / name
            outs = sorted(rdir.glob("result_CPU0_TIME*.out"), key=lambda p: int(p.stem.rsplit("TIME", 1)[1]))
            out = HERE / "vtk" / name
            shutil.copy(outs[-1], out / "final.out")
            shutil.copy(rdir / "visualisation.gnu", out / "visualisation.gnu")
            print(f"{name}: {outs[-1].name}, visualisation.gnu -> {out}")
        finally:
            shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main(sys.argv)
