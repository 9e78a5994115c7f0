"""Input decks of the reference ECOGEN executable -- TEST INFRASTRUCTURE ONLY.

Writes a run directory (ECOGEN.xml, libEOS/, case/*.xml) in the reference's
input schema (libTests/manual*.xml) for a case description with geometric
domains and states.  Used by tests/golden/make_golden.py (golden vectors)
and by the CPU-baseline leg of bench.py (the reference timed on the host).
EOS files carry the parameters of libEOS/SG_water.xml and libEOS/IG_air.xml.
"""
from __future__ import annotations

from pathlib import Path

EOS_XML = {
    "IG": '<?xml version="1.0" encoding="UTF-8" standalone="yes"?>\n<parametersEOS>\n'
          '  <EOS type="IG"/>\n  <parameters gamma="{gamma!r}" cv="{cv!r}" energyRef="{eref!r}" entropyRef="{sref!r}"/>\n'
          '</parametersEOS>\n',
    "SG": '<?xml version="1.0" encoding="UTF-8" standalone="yes"?>\n<parametersEOS>\n'
          '  <EOS type="SG"/>\n  <parameters gamma="{gamma!r}" pInf="{pinf!r}" cv="{cv!r}" energyRef="{eref!r}" entropyRef="{sref!r}"/>\n'
          '</parametersEOS>\n',
}


# full-precision ASCII output (golden vectors); binary output is Float32
OUTPUT_ASCII17 = 'format="XML" binary="false" precision="17"'
OUTPUT_BINARY = 'format="XML" binary="true"'


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
            if model in ("thermaleq", "eulerhomogeneous"):
                # PhaseThermalEq.cpp:45-57 and PhaseEulerHomogeneous.cpp:14-24 read alpha only
                out.append(f'      <dataFluid alpha="{st["alpha"][k]!r}"/>')
            elif model == "kapila":
                out.append(f'      <dataFluid alpha="{st["alpha"][k]!r}" density="{st["rho"][k]!r}"/>')
            else:
                out.append(f'      <dataFluid alpha="{st["alpha"][k]!r}" density="{st["rho"][k]!r}" pressure="{st["p"][k]!r}"/>')
            out.append('    </material>')
        u = list(st["u"]) + [0.0] * (3 - len(st["u"]))
        out.append('    <mixture>')
        if model in ("kapila", "eulerhomogeneous"):
            # MixKapila, MixEulerHomogeneous.cpp:25-43: the mixture pressure
            out.append(f'      <dataMix pressure="{st["p"][0]!r}"/>')
        elif model == "thermaleq":
            # MixThermalEq.cpp:46-79: pressure and temperature of the mixture
            out.append(f'      <dataMix pressure="{st["p"][0]!r}" temperature="{st["T"]!r}"/>')
        out.append(f'      <velocity x="{u[0]!r}" y="{u[1]!r}" z="{u[2]!r}"/>')
        out.append('    </mixture>')
    # Input.cpp:771-790: <transport name value> of the state, matched by name
    for name, v in zip(case.get("transports", []), st.get("tr", [])):
        out.append(f'    <transport name="{name}" value="{v!r}"/>')
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
        (root / "libEOS" / _eos_file(e)).write_text(EOS_XML[e["type"]].format(**{"sref": 0.0, "pinf": 0.0, **e}))
    (root / "ECOGEN.xml").write_text(
        '<?xml version="1.0" encoding="UTF-8" standalone="yes"?>\n<ecogen>\n'
        '  <testCase>./case/</testCase>\n</ecogen>\n')
    if case["time_mode"] == "iterations":
        freq = case.get("iter_freq", case["steps"])
        tc = (f'  <timeControlMode iterations="true">\n    <iterations number="{case["steps"]}" iterFreq="{freq}"/>\n'
              f'    <physicalTime totalTime="1." timeFreq="1."/>\n  </timeControlMode>')
    else:
        tf = case["t_final"]
        tfq = case.get("time_freq", tf)
        tc = (f'  <timeControlMode iterations="false">\n    <iterations number="1" iterFreq="1"/>\n'
              f'    <physicalTime totalTime="{tf!r}" timeFreq="{tfq!r}"/>\n  </timeControlMode>')
    for pr in case.get("probes", []):
        # Input.cpp:136-142, OutputProbeGNU.cpp:44-90: <probe name> with vertex and acqFreq
        v = pr["vertex"]
        tc += (f'\n  <probe name="{pr["name"]}">\n    <vertex x="{v[0]!r}" y="{v[1]!r}" z="{v[2]!r}"/>\n'
               f'    <timeControl acqFreq="{pr["acq_freq"]!r}"/>\n  </probe>')
    for cu in case.get("cuts", []):
        # Input.cpp:122-135, OutputCutGNU.cpp:44-90: <cut1D> with vertex and
        # vecDir, <cut2D> with vertex and vecNormal
        v, w = cu["vertex"], cu["vec"]
        tag, vt = ("cut1D", "vecDir") if cu["type"] == "cut1D" else ("cut2D", "vecNormal")
        tc += (f'\n  <{tag}>\n    <vertex x="{v[0]!r}" y="{v[1]!r}" z="{v[2]!r}"/>\n'
               f'    <{vt} x="{w[0]!r}" y="{w[1]!r}" z="{w[2]!r}"/>\n  </{tag}>')
    if case.get("resume"):
        # Input.cpp:253-257: restart from result file number fileNumber
        tc += f'\n  <resumeSimulation fileNumber="{case["resume"]}"/>'
    
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
        f'  <run>{run}</run>\n  <outputMode {case.get("output", OUTPUT_ASCII17)}/>\n'
        f'{tc}\n  <computationControl CFL="{case["cfl"]!r}"/>\n{so}</computationParam>\n')
    L = case["lengths"]
    n = case["cells"]
    (cdir / "meshV5.xml").write_text(
        '<?xml version="1.0" encoding="UTF-8" standalone="yes"?>\n<mesh>\n  <type structure="cartesian"/>\n'
        f'  <cartesianMesh>\n    <dimensions x="{L[0]!r}" y="{L[1]!r}" z="{L[2]!r}"/>\n'
        f'    <numberCells x="{n[0]}" y="{n[1]}" z="{n[2]}"/>\n  </cartesianMesh>\n</mesh>\n')
    model = case["model"]
    fm = {"euler": '<flowModel name="Euler"/>',
          "kapila": f'<flowModel name="Kapila" numberPhases="{len(case["eos"])}" alphaNull="{"true" if case.get("alpha_null") else "false"}"/>',
          "multip": f'<flowModel name="MultiP" numberPhases="{len(case["eos"])}"/>',
          "thermaleq": f'<flowModel name="ThermalEq" numberPhases="{len(case["eos"])}"/>',
          # Input.cpp:479-483: the liquid and vapour phase indices
          "eulerhomogeneous": f'<flowModel name="EulerHomogeneous" liquid="{case.get("liq", 0)}" '
                              f'vapor="{case.get("vap", 1)}"/>'}[model]
    trs = case.get("transports", [])
    if trs:
        # Input.cpp:472-473, 528-540: numberTransports and one <transport name> each
        fm = fm.replace("/>", f' numberTransports="{len(trs)}"/>')
    eosl = "\n".join(f'  <EOS name="{_eos_file(e)}"/>' for e in case["eos"])
    eosl += "".join(f'\n  <transport name="{t}"/>' for t in trs)
    relax = '  <relaxation type="P"/>\n' if case.get("relax_p") and model == "multip" else ""
    if case.get("relax_pt"):
        # Input.cpp:607-621: relaxations in deck order, after Kapila's own P
        relax += '  <relaxation type="PT"/>\n'
    if case.get("relax_ptmu"):
        # RelaxationPTMu.cpp:37-49 requires <dataPTMu> and takes phase 0 as the
        # liquid, phase 1 as the vapour whatever its attributes say
        e = case["eos"]
        relax += (f'  <relaxation type="PTMu">\n    <dataPTMu liquid="{_eos_file(e[0])}" vapor="{_eos_file(e[1])}"/>\n'
                  '  </relaxation>\n')
    (cdir / "modelV4.xml").write_text(
        '<?xml version="1.0" encoding="UTF-8" standalone="yes"?>\n<model>\n'
        f'  {fm}\n{eosl}\n{relax}</model>\n')
    doms = "\n".join(_domain_xml(i, d) for i, d in enumerate(case["domains"]))
    names = ["CLXm", "CLXp", "CLYm", "CLYp", "CLZm", "CLZp"]
    dim = sum(1 for c in n if c != 1)
    kinds = case.get("bc", ["abs"] * 6)
    p0s = case.get("bc_p0", [0.0] * 6)

    def tr_xml(i, tag, attrs):
        # BoundCondOutflow.cpp:58-79, BoundCondInj.cpp:120-147, BoundCondTank.cpp:146-173:
        # the side's transport values inside its data element
        vals = (case.get("bc_tr") or [None] * 6)[i] or []
        if not trs:
            return f'      <{tag} {attrs}/>\n'
        body = "".join(f'        <transport name="{t}" value="{v!r}"/>\n' for t, v in zip(trs, vals))
        return f'      <{tag} {attrs}>\n{body}      </{tag}>\n'

    def bc_xml(i):
        head = f'    <boundCond name="{names[i]}" type="{kinds[i]}" number="{i + 1}"'
        if kinds[i] == "outflow":
            return head + '>\n' + tr_xml(i, "dataOutflow", f'p0="{p0s[i]!r}"') + '    </boundCond>'
        if kinds[i] == "injection":
            inj = case["bc_inj"][i]
            body = '>\n' + tr_xml(i, "dataInjection", f'm0="{inj["m0"]!r}"')
            if len(case["eos"]) == 1:
                body += f'      <dataFluid density="{inj["rho"][0]!r}" pressure="{inj["p"][0]!r}"/>\n'
            else:
                body += "      <fluidsProp>\n" + "".join(
                    f'        <dataFluid EOS="{_eos_file(e)}" alpha="{inj["alpha"][k]!r}" density="{inj["rho"][k]!r}" '
                    f'pressure="{inj["p"][k]!r}"/>\n' for k, e in enumerate(case["eos"])) + "      </fluidsProp>\n"
            return head + body + "    </boundCond>"
        if kinds[i] == "tank":
            tk = case["bc_tank"][i]
            body = '>\n' + tr_xml(i, "dataTank", f'p0="{tk["p0"]!r}" T0="{tk["T0"]!r}"')
            if len(case["eos"]) > 1:
                key = "massFrac" if "massFrac" in tk else "alpha"
                body += "      <fluidsProp>\n" + "".join(
                    f'        <dataFluid EOS="{_eos_file(e)}" {key}="{tk[key][k]!r}"/>\n'
                    for k, e in enumerate(case["eos"])) + "      </fluidsProp>\n"
            return head + body + "    </boundCond>"
        return head + "/>"
    bcs = "\n".join(bc_xml(i) for i in range(2 * dim))
    states = "\n".join(_state_xml(case, st) for st in case["states"])
    (cdir / "initialConditionsV4.xml").write_text(
        '<?xml version="1.0" encoding="UTF-8" standalone="yes"?>\n<CI>\n  <physicalDomains>\n'
        f'{doms}\n  </physicalDomains>\n  <boundaryConditions>\n{bcs}\n  </boundaryConditions>\n{states}\n</CI>\n')
