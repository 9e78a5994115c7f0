"""Reference probe files for the output-writer tests.

Runs the reference executable (``oracle/_ref/ECOGEN``, built by
``oracle/ref.mk``) on small decks with <probe> elements (``OutputProbeGNU``)
and keeps, under ``tests/golden/probes/<case>/``, the probe files and their
gnuplot scripts, with the case description.  The files are the reference's
outputs: data.

    python tests/golden/make_probe_fixtures.py
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
from cases import _euler_1d, _shock_cylinders_2d  # noqa: E402
from deck import write_deck  # noqa: E402

REF_BIN = REPO / "oracle" / "_ref" / "ECOGEN"

FIXTURES = {
    "kapila_2d_fo_probes": dict(
        _shock_cylinders_2d(n=40, steps=20), order=1, limiters=["none"] * 4, iter_freq=10,
        probes=[{"name": "probeA", "vertex": [0.3, 0.5, 0.0], "acq_freq": 2e-6},
                {"name": "probeB", "vertex": [0.05, 0.05, 0.0], "acq_freq": 5e-6}]),
    "euler_1d_fo_probes": dict(
        _euler_1d(1, ["none"] * 4, 40, cells=50), iter_freq=20,
        probes=[{"name": "probeFace", "vertex": [0.4, 0.0, 0.0], "acq_freq": 1e-5},
                {"name": "probeEnd", "vertex": [0.99, 0.05, 0.05], "acq_freq": 3e-5}]),
}


def main():
    for name, case in FIXTURES.items():
        tmp = Path(tempfile.mkdtemp(prefix="ecogen_probes_"))
        try:
            write_deck(dict(case, name=name), tmp)
            r = subprocess.run([str(REF_BIN)], cwd=tmp, capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
            rdir = tmp / "results" / name
            out = HERE / "probes" / name
            if out.exists():
                shutil.rmtree(out)
            out.mkdir(parents=True)
            for pr in case.get("probes", []):
                shutil.copy(rdir / "probes" / f"{pr['name']}.out", out / f"{pr['name']}.out")
                shutil.copy(rdir / "probes" / f"visualisation{pr['name']}.gnu", out / f"visualisation{pr['name']}.gnu")
            (out / "case.json").write_text(json.dumps(case))
            print(f"{name}: {sorted(p.name for p in out.iterdir())}")
        finally:
            shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()
