"""Reference cut files for the output-writer tests.

Runs the reference executable (``oracle/_ref/ECOGEN``, built by
``oracle/ref.mk``) on small decks with <cut1D> / <cut2D> elements
(``OutputCutGNU``) and keeps, under ``tests/golden/cuts/<case>/``, every cut
file and gnuplot script of the run, with the case description.  The files are
the reference's outputs: data.

    python tests/golden/make_cut_fixtures.py
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
from cases import _bubbles_3d, _euler_1d, _shock_cylinders_2d  # noqa: E402
from deck import write_deck  # noqa: E402

REF_BIN = REPO / "oracle" / "_ref" / "ECOGEN"

FIXTURES = {
    # a 2.5 m box: the line's cells beyond 1 m along it are kept by distance
    # only; two cut1D share one file name (the later one remains)
    "kapila_2d_fo_cuts": dict(
        _shock_cylinders_2d(n=30, steps=4), lengths=[2.5, 2.5, 1.0], order=1, limiters=["none"] * 4, iter_freq=2,
        cuts=[{"type": "cut1D", "vertex": [0.0, 0.5, 0.0], "vec": [1.0, 0.0, 0.0]},
              {"type": "cut1D", "vertex": [0.3, 0.2, 0.0], "vec": [2.0, 1.0, 0.0]},
              {"type": "cut2D", "vertex": [1.3, 0.5, 0.0], "vec": [1.0, 0.2, 0.0]}]),
    "euler_1d_fo_cuts": dict(
        _euler_1d(1, ["none"] * 4, 6, cells=60), lengths=[3.0, 0.1, 0.1], iter_freq=3,
        cuts=[{"type": "cut1D", "vertex": [0.7, 0.05, 0.05], "vec": [1.0, 0.0, 0.0]}]),
    # Kapila: the MultiP + relaxation files print temperatures cached before
    # the last relaxation, which the stored state does not hold
    "kapila_3d_o2_cuts": dict(
        _bubbles_3d(n=10, steps=2, model="kapila"), iter_freq=2,
        cuts=[{"type": "cut1D", "vertex": [0.5, 0.5, 0.5], "vec": [0.0, 0.0, 1.0]},
              {"type": "cut2D", "vertex": [0.5, 0.45, 0.5], "vec": [0.0, 1.0, 0.0]}]),
}


def main():
    for name, case in FIXTURES.items():
        tmp = Path(tempfile.mkdtemp(prefix="ecogen_cuts_"))
        try:
            write_deck(dict(case, name=name), tmp)
            r = subprocess.run([str(REF_BIN)], cwd=tmp, capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
            src = tmp / "results" / name / "cuts"
            out = HERE / "cuts" / name
            if out.exists():
                shutil.rmtree(out)
            out.mkdir(parents=True)
            for f in sorted(src.iterdir()):
                shutil.copy(f, out / f.name)
            (out / "case.json").write_text(json.dumps(case))
            print(f"{name}: {sorted(p.name for p in out.iterdir())}")
        finally:
            shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()
