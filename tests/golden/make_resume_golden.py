"""Write golden vectors for resumed runs (<resumeSimulation fileNumber="N"/>)
from the reference.

For each case the upstream executable (``oracle/_ref/ECOGEN``, built from
source by ``oracle/ref.mk``) runs the deck twice in one run directory: first
uninterrupted with intermediate result files, then again with
``resumeSimulation`` set, so that it reads back ``infoCalcul.out`` and result
file N of the first run (``Run::resumeSimulation``, src/Run.cpp:164-190) and
marches to the end.  Stored per case (data only):

* ``resume_vtr`` / ``info_run1``: the text of result file N and of
  ``infoCalcul.out`` as the first run wrote them (what a resumed run reads);
* ``final_*``: the resumed run's final state; ``info``: its ``infoCalcul.out``
  rows (iteration, time, dt), truncated after row N and continued;
* ``run1_final_*``: the uninterrupted run's final state.

Usage (in the development container, where the reference tree exists):

    make -f oracle/ref.mk -j8
    python tests/golden/make_resume_golden.py
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
sys.path.insert(0, str(HERE))
from make_golden import REF_BIN, fields_from_vtr, parse_vtr, write_deck  # noqa: E402
from resume_cases import RESUME_CASES  # noqa: E402


def _rows(path):
    return np.array([[float(x) for x in r.split()] for r in path.read_text().split("\n")[1:] if r.strip()])


def run_resume(name, case):
    tmp = Path(tempfile.mkdtemp(prefix="ecogen_resume_"))
    try:
        first = dict(case, name=name)
        first.pop("resume")
        write_deck(first, tmp)
        for stage in ("run1", "run2"):
            if stage == "run2":
                shutil.rmtree(tmp / "case")
                shutil.rmtree(tmp / "libEOS")
                write_deck(dict(case, name=name), tmp)  # the results folder stays
            r = subprocess.run([str(REF_BIN)], cwd=tmp, capture_output=True, text=True, timeout=3600)
            if r.returncode != 0:
                raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
            rdir = tmp / "results" / name
            vtrs = sorted(rdir.glob("result_CPU0_AMR0_TIME*.vtr"), key=lambda p: int(p.stem.rsplit("TIME", 1)[1]))
            if stage == "run1":
                n = case["resume"]
                resume_vtr = (rdir / f"result_CPU0_AMR0_TIME{n}.vtr").read_text()
                info1 = (rdir / "infoCalcul.out").read_text()
                run1_final = fields_from_vtr(case, parse_vtr(vtrs[-1]))
                last1 = vtrs[-1].name
            else:
                if vtrs[-1].name != last1:
                    raise RuntimeError(f"resumed run ended at {vtrs[-1].name}, the first at {last1}")
                final = fields_from_vtr(case, parse_vtr(vtrs[-1]))
                info = _rows(rdir / "infoCalcul.out")
        return resume_vtr, info1, run1_final, final, info
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def main(argv):
    names = argv[1:] or list(RESUME_CASES)
    for name in names:
        case = RESUME_CASES[name]
        resume_vtr, info1, run1_final, final, info = run_resume(name, case)
        arrs = {f"final_{k}": v for k, v in final.items()}
        arrs.update({f"run1_final_{k}": v for k, v in run1_final.items()})
        arrs["info"] = info
        arrs["resume_vtr"] = np.frombuffer(resume_vtr.encode(), dtype=np.uint8)
        arrs["info_run1"] = np.frombuffer(info1.encode(), dtype=np.uint8)
        arrs["case_json"] = np.frombuffer(json.dumps(case).encode(), dtype=np.uint8)
        np.savez_compressed(HERE / "resume" / f"{name}.npz", **arrs)
        print(f"{name}: resumed from file {case['resume']} (it {int(info[case['resume'], 1])}), "
              f"ended at it {int(info[-1, 1])} t={info[-1, 2]!r}")


if __name__ == "__main__":
    main(sys.argv)
