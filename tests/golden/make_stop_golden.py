"""Write the fixtures of reference runs that stop on an error.

For each case of ``cases.STOP_CASES`` the reference executable (built by
``oracle/ref.mk``) runs the deck with a result file at every iteration; its
process ends with the error message of Errors::errorMessage.  The fixture
(``tests/golden/stop/<name>.npz``) keeps the case, the infoCalcul.out rows
(result number, iteration, time, computation time, dt: the last row is the
last completed step) and the message the reference printed.

    make -f oracle/ref.mk -j8
    python tests/golden/make_stop_golden.py
"""
from __future__ import annotations

import json
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(REPO / "oracle"))
from cases import STOP_CASES  # noqa: E402
from deck import write_deck  # noqa: E402

REF_BIN = REPO / "oracle" / "_ref" / "ECOGEN"
OUT = HERE / "stop"


def main(argv):
    OUT.mkdir(exist_ok=True)
    for name in argv[1:] or list(STOP_CASES):
        case = dict(STOP_CASES[name], name=name)
        tmp = Path(tempfile.mkdtemp(prefix="ecogen_stop_"))
        try:
            write_deck(case, tmp)
            r = subprocess.run([str(REF_BIN)], cwd=tmp, capture_output=True, text=True, timeout=3600)
            msg = [ln for ln in (r.stdout + r.stderr).split("\n") if "not converged" in ln]
            if not msg:
                raise RuntimeError(f"{name}: the reference did not stop on an error\n{r.stdout[-2000:]}")
            rows = (tmp / "results" / name / "infoCalcul.out").read_text().split("\n")[1:]
            info = np.array([[float(x) for x in row.split()] for row in rows if row.strip()])
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        np.savez_compressed(OUT / f"{name}.npz", info=info,
                            message=np.frombuffer(msg[0].strip().encode(), dtype=np.uint8),
                            case_json=np.frombuffer(json.dumps(STOP_CASES[name]).encode(), dtype=np.uint8))
        print(f"{name}: stopped after iteration {int(info[-1, 1])}: {msg[0].strip()!r}")


if __name__ == "__main__":
    main(sys.argv)
