"""Write golden vectors for the explicit finite-volume step from the reference.

Runs the upstream ECOGEN executable built from source by ``oracle/ref.mk``
(``oracle/_ref/ECOGEN``) on small decks written here, and stores, per case,
the t=0 state and the final state it prints (VTK XML, ASCII, 17 digits), plus
the (iteration, t, dt) log of ``infoCalcul.out``.  Fixtures are data only:
inputs and outputs, one compressed ``.npz`` per case.

Usage (in the development container, where the reference tree exists):

    make -f oracle/ref.mk -j8
    python tests/golden/make_golden.py            # all cases
    python tests/golden/make_golden.py c1_kapila_fo

Decks follow the reference input schema (``libTests/manual*.xml``); the EOS
files are written here from the same parameters as ``libEOS/SG_water.xml`` and
``libEOS/IG_air.xml`` (gamma, pInf, cv, energyRef, entropyRef).
"""
from __future__ import annotations

import json
import re
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
from cases import CASES  # noqa: E402
from deck import write_deck  # noqa: E402

REF_BIN = REPO / "oracle" / "_ref" / "ECOGEN"

_ARR = re.compile(r'<DataArray type="Float32" Name="([^"]+)"(?: NumberOfComponents="(\d)")? format="ascii">\s*(.*?)\s*</DataArray>', re.S)


def parse_vtr(path: Path):
    txt = path.read_text()
    cell = txt[txt.index("<CellData>"):txt.index("</CellData>")]
    out = {}
    for name, nc, body in _ARR.findall(cell):
        v = np.array(body.split(), dtype=np.float64)
        if nc:
            v = v.reshape(-1, int(nc))
        out[name] = v
    return out


def fields_from_vtr(case, d):
    """Map the printed arrays to (alpha_k, rho_k, p_k, u, v, w)."""
    model = case["model"]
    nph = len(case["eos"])
    res = {}
    if model in ("thermaleq", "eulerhomogeneous"):
        # PhaseThermalEq (PhaseEulerHomogeneous) prints Alpha and Density,
        # MixThermalEq (MixEulerHomogeneous) the mixture
        # density, pressure, temperature and velocity (ThermalEq*.cpp returnName*)
        res["alpha"] = np.stack([d[[k for k in d if k.startswith(f"F{i}_Alpha")][0]] for i in range(nph)])
        res["rho"] = np.stack([d[[k for k in d if k.startswith(f"F{i}_Density")][0]] for i in range(nph)])
        res["p"] = np.stack([d["Pressure_Mixture"]] * nph)
        res["p_mix"] = d["Pressure_Mixture"]
        res["rho_mix"] = d["Density_Mixture"]
        res["T"] = d["Temperature_Mixture"]
        res["u"] = (d["Velocity_Mixture_norm"] if model == "thermaleq" else d["Velocity_Mixture"]).T.copy()
        return res
    if model == "euler":
        key = [k for k in d if k.startswith("F0_Density")][0]
        res["rho"] = d[key][None, :]
        res["p"] = d[[k for k in d if k.startswith("F0_Pressure")][0]][None, :]
        vel = d[[k for k in d if k.startswith("F0_Velocity")][0]]
    else:
        res["alpha"] = np.stack([d[[k for k in d if k.startswith(f"F{i}_Alpha")][0]] for i in range(nph)])
        res["rho"] = np.stack([d[[k for k in d if k.startswith(f"F{i}_Density")][0]] for i in range(nph)])
        if model == "multip":
            res["p"] = np.stack([d[[k for k in d if k.startswith(f"F{i}_Pressure")][0]] for i in range(nph)])
        else:
            res["p"] = np.stack([d["Pressure_Mixture"]] * nph)
        res["p_mix"] = d["Pressure_Mixture"]
        res["rho_mix"] = d["Density_Mixture"]
        res["temp"] = np.stack([d[[k for k in d if k.startswith(f"F{i}_Temperature")][0]] for i in range(nph)])
        vel = d["Velocity_Mixture"]
    res["u"] = vel.T.copy()
    # transported scalars: DataArray "T1", "T2", ... (OutputXML.cpp:435-447)
    ntr = len(case.get("transports", []))
    if ntr:
        res["tr"] = np.stack([d[f"T{t + 1}"] for t in range(ntr)])
    return res


def run_case(case, keep=False):
    tmp = Path(tempfile.mkdtemp(prefix="ecogen_golden_"))
    try:
        write_deck(case, tmp)
        r = subprocess.run([str(REF_BIN)], cwd=tmp, capture_output=True, text=True, timeout=3600)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
        rdir = tmp / "results" / case["name"]
        vtrs = [p for p in rdir.glob("result_CPU0_AMR0_TIME*.vtr")]
        vtrs.sort(key=lambda p: int(p.stem.rsplit("TIME", 1)[1]))
        if len(vtrs) < 2:
            raise RuntimeError(f"expected >=2 result files, got {vtrs}\n{r.stdout[-3000:]}")
        init = fields_from_vtr(case, parse_vtr(vtrs[0]))
        final = fields_from_vtr(case, parse_vtr(vtrs[-1]))
        rows = (rdir / "infoCalcul.out").read_text().split("\n")[1:]
        info = np.array([[float(x) for x in r.split()] for r in rows if r.strip()])
        return init, final, info, r.stdout
    finally:
        if not keep:
            shutil.rmtree(tmp, ignore_errors=True)
        else:
            print("kept", tmp)


def main(argv):
    names = argv[1:] or list(CASES)
    for name in names:
        case = dict(CASES[name], name=name)
        init, final, info, _ = run_case(case)
        arrs = {f"init_{k}": v for k, v in init.items()}
        arrs.update({f"final_{k}": v for k, v in final.items()})
        arrs["info"] = info
        arrs["case_json"] = np.frombuffer(json.dumps(CASES[name]).encode(), dtype=np.uint8)
        np.savez_compressed(HERE / f"{name}.npz", **arrs)
        print(f"{name}: it={int(info[-1, 1])} t={info[-1, 2]!r} -> {name}.npz")


if __name__ == "__main__":
    main(sys.argv)
