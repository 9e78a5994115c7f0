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
from cases import CASES  # noqa: E402

REF_BIN = REPO / "oracle" / "_ref" / "ECOGEN"

EOS_XML = {
    "IG": '<?xml version="1.0" encoding="UTF-8" standalone="yes"?>\n<parametersEOS>\n'
          '  <EOS type="IG"/>\n  <parameters gamma="{gamma!r}" cv="{cv!r}" energyRef="{eref!r}" entropyRef="{sref!r}"/>\n'
          '</parametersEOS>\n',
    "SG": '<?xml version="1.0" encoding="UTF-8" standalone="yes"?>\n<parametersEOS>\n'
          '  <EOS type="SG"/>\n  <parameters gamma="{gamma!r}" pInf="{pinf!r}" cv="{cv!r}" energyRef="{eref!r}" entropyRef="{sref!r}"/>\n'
          '</parametersEOS>\n',
}


def _eos_file(e):
    return f"{e['type']}_{e['name']}.xml"


def _state_xml(case, st):
    model = case["model"]
    out = [f'  <state name="{st["name"]}">']
    eos = case["eos"]
    if model == "euler":
        e = eos[0]
        out.append(f'    <material type="fluide" EOS="{_eos_file(e)}">')
        out.append(f'      <dataFluid density="{st["rho"][0]!r}" pressure="{st["p"][0]!r}">')
        u = list(st["u"]) + [0.0] * (3 - len(st["u"]))
        out.append(f'        <velocity x="{u[0]!r}" y="{u[1]!r}" z="{u[2]!r}"/>')
        out.append('      </dataFluid>')
        out.append('    </material>')
    else:
        for k, e in enumerate(eos):
            out.append(f'    <material type="fluide" EOS="{_eos_file(e)}">')
            if model == "kapila":
                out.append(f'      <dataFluid alpha="{st["alpha"][k]!r}" density="{st["rho"][k]!r}"/>')
            else:
                out.append(f'      <dataFluid alpha="{st["alpha"][k]!r}" density="{st["rho"][k]!r}" pressure="{st["p"][k]!r}"/>')
            out.append('    </material>')
        u = list(st["u"]) + [0.0] * (3 - len(st["u"]))
        out.append('    <mixture>')
        if model == "kapila":
            out.append(f'      <dataMix pressure="{st["p"][0]!r}"/>')
        out.append(f'      <velocity x="{u[0]!r}" y="{u[1]!r}" z="{u[2]!r}"/>')
        out.append('    </mixture>')
    out.append('  </state>')
    return "\n".join(out)


def _domain_xml(i, d):
    t = d["type"]
    head = f'    <domain name="d{i}" state="{d["state"]}" type="{t}"'
    if t == "entireDomain":
        return head + "/>"
    if t == "halfSpace":
        return (head + ">\n" + f'      <dataHalfSpace axe="{d["axe"]}" origin="{d["origin"]!r}" direction="{d["direction"]}"/>\n'
                "    </domain>")
    if t == "disc":
        c = d["center"]
        return (head + ">\n" + f'      <dataDisc axe1="x" axe2="y" radius="{d["radius"]!r}">\n'
                f'        <center x="{c[0]!r}" y="{c[1]!r}" z="0."/>\n      </dataDisc>\n    </domain>')
    if t == "sphere":
        c = d["center"]
        return (head + ">\n" + f'      <dataSphere radius="{d["radius"]!r}">\n'
                f'        <center x="{c[0]!r}" y="{c[1]!r}" z="{c[2]!r}"/>\n      </dataSphere>\n    </domain>')
    raise ValueError(t)


def write_deck(case, root: Path):
    run = case["name"]
    cdir = root / "case"
    cdir.mkdir(parents=True)
    (root / "libEOS").mkdir()
    for e in case["eos"]:
        (root / "libEOS" / _eos_file(e)).write_text(EOS_XML[e["type"]].format(**e))
    (root / "ECOGEN.xml").write_text(
        '<?xml version="1.0" encoding="UTF-8" standalone="yes"?>\n<ecogen>\n'
        '  <testCase>./case/</testCase>\n</ecogen>\n')
    if case["time_mode"] == "iterations":
        tc = (f'  <timeControlMode iterations="true">\n    <iterations number="{case["steps"]}" iterFreq="{case["steps"]}"/>\n'
              f'    <physicalTime totalTime="1." timeFreq="1."/>\n  </timeControlMode>')
    else:
        tf = case["t_final"]
        tc = (f'  <timeControlMode iterations="false">\n    <iterations number="1" iterFreq="1"/>\n'
              f'    <physicalTime totalTime="{tf!r}" timeFreq="{tf!r}"/>\n  </timeControlMode>')
    so = ""
    if case["order"] == 2:
        lim = case["limiters"]
        so = ("  <secondOrder>\n"
              f"    <globalLimiter>{lim[0]}</globalLimiter>\n"
              f"    <interfaceLimiter>{lim[1]}</interfaceLimiter>\n"
              f"    <globalVolumeFractionLimiter>{lim[2]}</globalVolumeFractionLimiter>\n"
              f"    <interfaceVolumeFractionLimiter>{lim[3]}</interfaceVolumeFractionLimiter>\n"
              "  </secondOrder>\n")
    (cdir / "mainV5.xml").write_text(
        '<?xml version="1.0" encoding="UTF-8" standalone="yes"?>\n<computationParam>\n'
        f'  <run>{run}</run>\n  <outputMode format="XML" binary="false" precision="17"/>\n'
        f'{tc}\n  <computationControl CFL="{case["cfl"]!r}"/>\n{so}</computationParam>\n')
    L = case["lengths"]
    n = case["cells"]
    (cdir / "meshV5.xml").write_text(
        '<?xml version="1.0" encoding="UTF-8" standalone="yes"?>\n<mesh>\n  <type structure="cartesian"/>\n'
        f'  <cartesianMesh>\n    <dimensions x="{L[0]!r}" y="{L[1]!r}" z="{L[2]!r}"/>\n'
        f'    <numberCells x="{n[0]}" y="{n[1]}" z="{n[2]}"/>\n  </cartesianMesh>\n</mesh>\n')
    model = case["model"]
    fm = {"euler": '<flowModel name="Euler"/>',
          "kapila": f'<flowModel name="Kapila" numberPhases="{len(case["eos"])}" alphaNull="{"true" if case["alpha_null"] else "false"}"/>',
          "multip": f'<flowModel name="MultiP" numberPhases="{len(case["eos"])}"/>'}[model]
    eosl = "\n".join(f'  <EOS name="{_eos_file(e)}"/>' for e in case["eos"])
    relax = '  <relaxation type="P"/>\n' if case.get("relax_p") and model == "multip" else ""
    (cdir / "modelV4.xml").write_text(
        '<?xml version="1.0" encoding="UTF-8" standalone="yes"?>\n<model>\n'
        f'  {fm}\n{eosl}\n{relax}</model>\n')
    doms = "\n".join(_domain_xml(i, d) for i, d in enumerate(case["domains"]))
    names = ["CLXm", "CLXp", "CLYm", "CLYp", "CLZm", "CLZp"]
    dim = sum(1 for c in n if c != 1)
    bcs = "\n".join(f'    <boundCond name="{names[i]}" type="abs" number="{i + 1}"/>' for i in range(2 * dim))
    states = "\n".join(_state_xml(case, st) for st in case["states"])
    (cdir / "initialConditionsV4.xml").write_text(
        '<?xml version="1.0" encoding="UTF-8" standalone="yes"?>\n<CI>\n  <physicalDomains>\n'
        f'{doms}\n  </physicalDomains>\n  <boundaryConditions>\n{bcs}\n  </boundaryConditions>\n{states}\n</CI>\n')


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
