"""Reader of the reference's input decks, so that an ECOGEN run directory runs
unchanged on the device path.

A run directory holds ``ECOGEN.xml`` (the test-case folder), the case files
``mainV5.xml``, ``meshV5.xml``, ``modelV4.xml`` and ``initialConditionsV4.xml``,
and the EOS files under ``libEOS/``.  ``load`` turns it into the case
description of :class:`ecogen_amd.Context` and the initial state in the
``ecogen.h`` host layout, following ``Input::entreeMain``, ``entreeMesh``,
``entreeModel``, ``entreeEOS`` and ``entreeConditionsInitiales``
(``src/InputOutput/Input.cpp:85-864``) for what the device path runs:

* Cartesian meshes without AMR;
* the Euler, Kapila and MultiP flow models, ideal- and stiffened-gas EOS, the
  MultiP pressure relaxation;
* first order, or MUSCL with the four limiters and their defaults
  (``Input.cpp:185-249``);
* iteration or physical-time control and the CFL number;
* the geometric domains entireDomain, halfSpace, disc, sphere and rectangle,
  applied in deck order (``Cell::fill``, ``src/Cell.cpp:130-137``; the
  ``belong`` tests of ``src/Geometries/GD*.cpp``);
* ``abs``, ``outflow`` (``BoundCondOutflow``, the exit pressure ``p0``),
  ``injection`` (``BoundCondInj``, the mass flow ``m0`` and the injected fluid),
  ``tank`` (``BoundCondTank``, a reservoir at ``p0``, ``T0``),
  ``wall`` and ``symmetry`` boundaries: ``BoundCondWall`` at first
  order for both wall and symmetry (``Input.cpp:849-852``), ``BoundCondWallO2``
  and ``BoundCondSymmetryO2`` at second order, which differ only in name
  (``src/BoundConds/BoundCondSymmetryO2.h:41``).

* ``<resumeSimulation fileNumber="N"/>``: the run restarts from result file N
  and row N of ``infoCalcul.out`` under ``RUNDIR/results/<run>/``, with the
  dt printed there for its first step (``Run::resumeSimulation``,
  ``src/Run.cpp:164-190``); ``--vtk`` continues the files and ``infoCalcul.out``.

Anything else raises ``DeckError`` naming the element.

Command line (runs the deck on the GPU and saves the final state; ``--vtk``
also writes the reference's VTK result files at the deck's output frequency,
``ecogen_amd/output.py``)::

    python -m ecogen_amd.deck RUNDIR [--out final.npz] [--precision fp64|fp32s] [--vtk]
"""
from __future__ import annotations

import argparse
import json
import math
import re
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np

LIMITERS = {"MINMOD": "minmod", "VANLEER": "vanleer", "VANALBADA": "vanalbada", "SUPERBEE": "superbee",
            "MC": "mc", "THINC": "thinc"}
MODELS = {"EULER": "euler", "KAPILA": "kapila", "MULTIP": "multip", "THERMALEQ": "thermaleq",
          "EULERHOMOGENEOUS": "eulerhomogeneous"}
AXES = {"X": 0, "Y": 1, "Z": 2}


class DeckError(ValueError):
    """The deck uses something the device path does not run, or is malformed."""


def _xml(path: Path):
    if not path.exists():
        raise DeckError(f"missing {path}")
    return ET.parse(path).getroot()


def _child(node, tag, where):
    c = node.find(tag)
    if c is None:
        raise DeckError(f"{where}: missing <{tag}>")
    return c


_NUM = re.compile(r"\s*[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?")


def _num(v, where=""):
    """A number as tinyxml2's QueryDoubleAttribute reads it (sscanf %lf): the
    longest numeric prefix, so that libEOS's energyRef="-1.15e6." is -1.15e6."""
    m = _NUM.match(v)
    if not m:
        raise DeckError(f"{where}: {v!r} is not a number")
    return float(m.group(0))


def _f(node, attr, where):
    v = node.get(attr)
    if v is None:
        raise DeckError(f"{where}: <{node.tag}> lacks {attr}")
    return _num(v, where)


def _true(v):
    return v is not None and v.strip().lower() == "true"


def _eos(path: Path):
    """libEOS file: EosIG / EosSG parameters (Eos*::assignParametersFromXML)."""
    root = _xml(path)
    kind = _child(root, "EOS", path.name).get("type", "").upper()
    par = _child(root, "parameters", path.name)
    # energyRef, entropyRef (the saturation temperature's reference entropies,
    # Mixture::computeTsat) default to 0
    eref = _num(par.get("energyRef", "0"), path.name)
    sref = _num(par.get("entropyRef", "0"), path.name)
    if kind == "IG":
        return {"type": "IG", "gamma": _f(par, "gamma", path.name), "pinf": 0.0, "cv": _f(par, "cv", path.name),
                "eref": eref, "sref": sref}
    if kind == "SG":
        return {"type": "SG", "gamma": _f(par, "gamma", path.name), "pinf": _f(par, "pInf", path.name),
                "cv": _f(par, "cv", path.name), "eref": eref, "sref": sref}
    raise DeckError(f"{path.name}: EOS type {kind!r} is not IG or SG")


def _limiters(so):
    """Input.cpp:185-249: interface defaults to global, global volume fraction
    to global, interface volume fraction to interface."""
    def text(tag):
        e = so.find(tag)
        if e is None:
            return None
        if e.text is None:
            raise DeckError(f"secondOrder: empty <{tag}>")
        return e.text.strip().upper()
    g = text("globalLimiter")
    if g is None:
        raise DeckError("secondOrder: missing <globalLimiter>")
    i = text("interfaceLimiter") or g
    gv = text("globalVolumeFractionLimiter") or g
    iv = text("interfaceVolumeFractionLimiter") or i
    for name, v in (("globalLimiter", g), ("interfaceLimiter", i)):
        if v == "THINC":
            raise DeckError(f"{name}: THINC can only be a volume-fraction limiter")
    out = []
    for v in (g, i, gv, iv):
        if v not in LIMITERS:
            raise DeckError(f"secondOrder: unknown limiter {v!r}")
        out.append(LIMITERS[v])
    return out


def _centres(cells, lengths):
    """Cell centres (i + 0.5) * L / n per axis (MeshCartesian)."""
    return [(np.arange(cells[d]) + 0.5) * (lengths[d] / cells[d]) for d in range(3)]


def _domain_mask(dom, X, Y, Z, where):
    """GeometricalDomain::belong of one domain on the cell-centre grids."""
    kind = dom.get("type", "")
    P = (X, Y, Z)
    if kind == "entireDomain":
        return np.ones(np.broadcast(X, Y, Z).shape, dtype=bool)
    if kind == "halfSpace":  # GDHalfSpace.cpp:74-96
        d = _child(dom, "dataHalfSpace", where)
        ax = AXES[d.get("axe", "").upper()]
        org = _f(d, "origin", where)
        direction = d.get("direction", "").lower()
        if direction not in ("positive", "negative"):
            raise DeckError(f"{where}: halfSpace direction {direction!r}")
        p = P[ax]
        m = (p >= org) if direction == "positive" else (p <= org)
        return np.broadcast_to(m, np.broadcast(X, Y, Z).shape)
    if kind == "disc":  # GDDisc.cpp:85-107
        d = _child(dom, "dataDisc", where)
        c = _child(d, "center", where)
        cen = [float(c.get(a, 0.0)) for a in "xyz"]
        r = _f(d, "radius", where)
        s = 0.0
        for a in (d.get("axe1", ""), d.get("axe2", "")):
            k = AXES[a.upper()]
            s = s + (P[k] - cen[k]) ** 2
        return np.broadcast_to(s <= r * r, np.broadcast(X, Y, Z).shape)
    if kind == "sphere":  # GDSphere.cpp:71-79
        d = _child(dom, "dataSphere", where)
        c = _child(d, "center", where)
        cen = [float(c.get(a, 0.0)) for a in "xyz"]
        r = _f(d, "radius", where)
        s = (X - cen[0]) ** 2 + (Y - cen[1]) ** 2 + (Z - cen[2]) ** 2
        return s <= r * r
    if kind == "rectangle":  # GDRectangle.cpp:88-120
        d = _child(dom, "dataRectangle", where)
        lb = _child(d, "posInferiorVertex", where)
        lo = [float(lb.get(a, 0.0)) for a in "xyz"]
        m = np.ones(np.broadcast(X, Y, Z).shape, dtype=bool)
        for a, la in ((d.get("axe1", ""), "lAxe1"), (d.get("axe2", ""), "lAxe2")):
            k = AXES[a.upper()]
            off = P[k] - lo[k]
            m = m & (off >= 0) & (off <= _f(d, la, where))
        return m
    raise DeckError(f"{where}: geometric domain type {kind!r} is not read")


def _named_values(node, names, where, need_all):
    """Values of the <transport name value> children of ``node`` matched by
    name against the deck's transports (Input.cpp:771-790 for a state, where
    a scalar not given is 0; BoundCondOutflow.cpp:58-79, BoundCondInj.cpp:
    120-147, BoundCondTank.cpp:146-173 for a boundary, which needs them all)."""
    vals = [0.0] * len(names)
    found = 0
    for tr in node.findall("transport") if node is not None else []:
        if tr.get("name") in names:
            vals[names.index(tr.get("name"))] = _f(tr, "value", where)
            found += 1
    if need_all and found < len(names):
        raise DeckError(f"{where}: not enough transport equations")
    return vals


def _tsat(eos, liq, vap, p):
    """Mixture::computeTsat (Mixture.cpp:63-104): the saturation temperature
    at pressure p, Newton on the equality of the phases' Gibbs energies, with
    the C library's log (math.log) in the reference's order of operations."""
    qL, qV = eos[liq], eos[vap]
    gL, pL, cvL, e0L, s0L = qL["gamma"], qL["pinf"], qL["cv"], qL["eref"], qL.get("sref", 0.0)
    gV, pV, cvV, e0V, s0V = qV["gamma"], qV["pinf"], qV["cv"], qV["eref"], qV.get("sref", 0.0)
    A = (gL * cvL - gV * cvV + s0V - s0L) / (gV * cvV - cvV)
    B = (e0L - e0V) / (gV * cvV - cvV)
    C = (gV * cvV - gL * cvL) / (gV * cvV - cvV)
    D = (gL * cvL - cvL) / (gV * cvV - cvV)
    T, f, df = 0.1 * B / C, 0.0, 1.0
    for _ in range(50):
        T -= f / df
        f = A + B / T + C * math.log(T) - math.log(p + pV) + D * math.log(p + pL)
        df = C / T - B / (T * T)
        if not abs(f) > 1e-10:
            break
    return T


def _state_values(st, model, nph, where, eos=None, liqvap=(0, 1)):
    """(alpha_k, rho_k, p_k per phase, u, v, w) of one <state>; for the
    thermal-equilibrium model also E (0, not read) and the temperature."""
    mats = st.findall("material")
    if len(mats) != nph:
        raise DeckError(f"{where}: {len(mats)} materials for {nph} phases")
    if model == "thermaleq":
        # PhaseThermalEq reads alpha (PhaseThermalEq.cpp:45-57), MixThermalEq the
        # mixture pressure, temperature and velocity (MixThermalEq.cpp:46-79);
        # rho_k = rho_k(p, T) as Eos::computeDensity (EosSG.cpp:114-117)
        mix = _child(st, "mixture", where)
        dm = _child(mix, "dataMix", where)
        p, T = _f(dm, "pressure", where), _f(dm, "temperature", where)
        vel = _child(mix, "velocity", where)
        out = []
        for m, e in zip(mats, eos):
            a = _f(_child(m, "dataFluid", where), "alpha", where)
            num = p + e["pinf"] if e["type"] == "SG" else p
            den = (e["gamma"] - 1.0) * e["cv"] * T
            out += [a, num / (0.0 if den < 0.0 else den), p]
        return out + [float(vel.get(a, 0.0)) for a in "xyz"] + [0.0, T]
    if model == "eulerhomogeneous":
        # PhaseEulerHomogeneous reads alpha (PhaseEulerHomogeneous.cpp:14-24),
        # MixEulerHomogeneous the mixture pressure and velocity
        # (MixEulerHomogeneous.cpp:25-43); the phases take the saturated
        # densities of p at the saturation temperature (ModEulerHomogeneous::
        # fulfillState, EosSG/EosIG::computeDensitySaturation)
        mix = _child(st, "mixture", where)
        p = _f(_child(mix, "dataMix", where), "pressure", where)
        vel = _child(mix, "velocity", where)
        T = _tsat(eos, liqvap[0], liqvap[1], p)
        out = []
        for m, e in zip(mats, eos):
            num = p + e["pinf"] if e["type"] == "SG" else p
            den = (e["gamma"] - 1.0) * e["cv"] * T
            out += [_f(_child(m, "dataFluid", where), "alpha", where), num / (0.0 if den < 0.0 else den), p]
        return out + [_num(vel.get(a, "0"), where) for a in "xyz"] + [0.0, T]
    if model == "euler":
        fl = _child(mats[0], "dataFluid", where)
        v = _child(fl, "velocity", where)
        return [1.0, _f(fl, "density", where), _f(fl, "pressure", where)] + [float(v.get(a, 0.0)) for a in "xyz"]
    mix = _child(st, "mixture", where)
    vel = _child(mix, "velocity", where)
    out = []
    pm = _f(_child(mix, "dataMix", where), "pressure", where) if model == "kapila" else None
    for m in mats:
        fl = _child(m, "dataFluid", where)
        p = pm if model == "kapila" else _f(fl, "pressure", where)
        out += [_f(fl, "alpha", where), _f(fl, "density", where), p]
    return out + [float(vel.get(a, 0.0)) for a in "xyz"]


def load(rundir):
    """(case, state) of a reference run directory.  ``state`` is the t = 0
    field-major array of the ecogen.h host layout without E."""
    rundir = Path(rundir)
    top = _xml(rundir / "ECOGEN.xml")
    tc = top.find("testCase")
    if tc is None or not (tc.text or "").strip():
        raise DeckError("ECOGEN.xml: missing <testCase>")
    cdir = (rundir / tc.text.strip()).resolve()
    main = _xml(cdir / "mainV5.xml")
    mesh = _xml(cdir / "meshV5.xml")
    model_x = _xml(cdir / "modelV4.xml")
    ci = _xml(cdir / "initialConditionsV4.xml")

    case = {"name": (main.findtext("run") or "deck").strip()}
    # run control (Input::entreeMain)
    tcm = _child(main, "timeControlMode", "mainV5.xml")
    if _true(tcm.get("iterations")):
        case["time_mode"] = "iterations"
        case["steps"] = int(_f(_child(tcm, "iterations", "mainV5.xml"), "number", "mainV5.xml"))
    else:
        case["time_mode"] = "time"
        case["t_final"] = _f(_child(tcm, "physicalTime", "mainV5.xml"), "totalTime", "mainV5.xml")
    case["cfl"] = _f(_child(main, "computationControl", "mainV5.xml"), "CFL", "mainV5.xml")
    # output cadence and format (Input::entreeMain; OutputXML)
    it_x, pt_x = tcm.find("iterations"), tcm.find("physicalTime")
    om = main.find("outputMode")
    case["output"] = {
        "iter_freq": int(float(it_x.get("iterFreq", "0"))) if it_x is not None else 0,
        "time_freq": float(pt_x.get("timeFreq", "0")) if pt_x is not None else 0.0,
        "format": (om.get("format") or "XML").upper() if om is not None else "XML",
        "binary": _true(om.get("binary")) if om is not None else False,
        "precision": int(float(om.get("precision", "0"))) if om is not None else 0,
    }
    # cuts (Input.cpp:122-135, OutputCutGNU.cpp:44-90): every <cut1D> (vertex,
    # vecDir), then every <cut2D> (vertex, vecNormal), in the reference's order
    case["cuts"] = []
    for tag, vt in (("cut1D", "vecDir"), ("cut2D", "vecNormal")):
        for cu in main.findall(tag):
            where = f"mainV5.xml {tag}"
            v, w = _child(cu, "vertex", where), _child(cu, vt, where)
            vec = [_f(w, a, where) for a in "xyz"]
            if math.sqrt(sum(c * c for c in vec)) < 1e-6:
                raise DeckError(f"{where}: null {vt}")  # GOLine/GOPlan constructors
            case["cuts"].append({"type": tag, "vertex": [_f(v, a, where) for a in "xyz"], "vec": vec})
    # probes (Input.cpp:136-142, OutputProbeGNU.cpp:44-90)
    case["probes"] = []
    for pr in main.findall("probe"):
        where = "mainV5.xml probe"
        v = _child(pr, "vertex", where)
        case["probes"].append({"name": pr.get("name") or "probe",
                               "vertex": [_f(v, a, where) for a in "xyz"],
                               "acq_freq": _f(_child(pr, "timeControl", where), "acqFreq", where)})
    rs = main.find("resumeSimulation")
    if rs is not None:
        # Input.cpp:253-257: restart from result file number fileNumber
        try:
            case["resume"] = int(rs.get("fileNumber", ""))
        except ValueError:
            raise DeckError("mainV5.xml: <resumeSimulation> lacks an integer fileNumber") from None
    if case["output"]["format"] not in ("XML", "GNU"):
        raise DeckError(f"mainV5.xml: output format {case['output']['format']!r} (XML or GNU)")
    if case["output"]["format"] == "GNU" and case["output"]["binary"]:
        raise DeckError("mainV5.xml: OutputGNU writes no binary data (OutputGNU.cpp:200-205)")
    so = main.find("secondOrder")
    if so is None:
        case["order"], case["limiters"] = 1, ["none"] * 4
    else:
        case["order"], case["limiters"] = 2, _limiters(so)

    # mesh (Input::entreeMesh)
    st_ = _child(mesh, "type", "meshV5.xml").get("structure", "").lower()
    if st_ != "cartesian":
        raise DeckError(f"meshV5.xml: mesh structure {st_!r} (only cartesian)")
    cm = _child(mesh, "cartesianMesh", "meshV5.xml")
    amr = cm.find("AMR")
    if amr is not None and int(float(amr.get("lvlMax", "0"))) > 0:
        raise DeckError("meshV5.xml: AMR stays on the reference")
    dims = _child(cm, "dimensions", "meshV5.xml")
    ncel = _child(cm, "numberCells", "meshV5.xml")
    case["lengths"] = [_f(dims, a, "meshV5.xml") for a in "xyz"]
    case["cells"] = [int(_f(ncel, a, "meshV5.xml")) for a in "xyz"]
    if case["output"]["format"] == "GNU" and all(c > 1 for c in case["cells"]):
        # OutputGNU::ecritScriptGnuplot stops the run at its first output in 3-D
        raise DeckError("mainV5.xml: GNU output in 3-D (OutputGNU.cpp:110)")

    # model and EOS (Input::entreeModel / entreeEOS)
    fm = _child(model_x, "flowModel", "modelV4.xml")
    name = fm.get("name", "").upper()
    if name not in MODELS:
        raise DeckError(f"modelV4.xml: flow model {fm.get('name')!r} is not Euler, Kapila or MultiP")
    model = MODELS[name]
    case["model"] = model
    case["alpha_null"] = _true(fm.get("alphaNull")) if model == "kapila" else False
    eos_files = [e.get("name", "") for e in model_x.findall("EOS")]
    if model == "eulerhomogeneous":
        # two phases, the liquid and vapour indices (Input.cpp:479-483)
        nph = 2
        case["liq"] = int(_f(fm, "liquid", "modelV4.xml"))
        case["vap"] = int(_f(fm, "vapor", "modelV4.xml"))
        if sorted((case["liq"], case["vap"])) != [0, 1]:
            raise DeckError("modelV4.xml: EulerHomogeneous takes liquid and vapor 0 and 1")
    else:
        nph = 1 if model == "euler" else int(_f(fm, "numberPhases", "modelV4.xml"))
    if len(eos_files) != nph:
        raise DeckError(f"modelV4.xml: {len(eos_files)} EOS for {nph} phases")
    case["eos"] = [dict(_eos(rundir / "libEOS" / f), name=Path(f).stem, file=f) for f in eos_files]
    # relaxations in deck order (Input.cpp:607-621), after Kapila's own P
    # relaxation (ModKapila.cpp:47-53): MultiP takes P, PT or P then PT,
    # Kapila PT; RelaxationPT needs two phases (RelaxationPT.cpp:60)
    relax = [r.get("type", "").upper() for r in model_x.findall("relaxation")]
    allowed = {"multip": (["P"], ["PT"], ["P", "PT"]), "kapila": (["PT"], ["PTMU"])}.get(model, ())
    if relax and relax not in [list(a) for a in allowed]:
        raise DeckError(f"modelV4.xml: relaxations {relax} for {model} (MultiP: P, PT or P then PT; Kapila: PT or PTMu)")
    if ("PT" in relax or "PTMU" in relax) and nph != 2:
        raise DeckError("modelV4.xml: the PT and PTMu relaxations take two phases")
    case["relax_p"] = model == "multip" and "P" in relax
    if "PT" in relax:
        case["relax_pt"] = True
    if "PTMU" in relax:
        # RelaxationPTMu.cpp:37-49: <dataPTMu> required; phase 0 is the liquid,
        # phase 1 the vapour whatever its attributes name
        if model_x.find("relaxation/dataPTMu") is None:
            raise DeckError("modelV4.xml: <relaxation type=\"PTMu\"> lacks <dataPTMu>")
        case["relax_ptmu"] = True
    # transport equations (Input.cpp:472-473, 528-540): numberTransports of
    # the first <transport name> elements
    ntr = int(float(fm.get("numberTransports", "0")))
    names = [t.get("name", "") for t in model_x.findall("transport")]
    if len(names) < ntr or any(not nm for nm in names[:ntr]):
        raise DeckError(f"modelV4.xml: {len(names)} named transports for numberTransports={ntr}")
    trs = names[:ntr]
    if trs:
        if model == "euler":
            # Model::solveRiemannTransportIntern ends the run (src/Models/Model.h:127)
            raise DeckError("modelV4.xml: the Euler model has no transport equations")
        case["transports"] = trs

    # initial conditions (Input::entreeConditionsInitiales)
    states = {s.get("name"): s for s in ci.findall("state")}
    bcs = ["abs"] * 6
    bc_tr = [None] * 6
    p0s = [0.0] * 6
    injs = [None] * 6
    tanks = [None] * 6
    bnode = ci.find("boundaryConditions")
    for b in (bnode.findall("boundCond") if bnode is not None else []):
        kind = b.get("type", "").lower()
        num = int(float(b.get("number", "0")))
        if not 1 <= num <= 6:
            raise DeckError(f"boundCond number {num}")
        if kind == "abs":
            bcs[num - 1] = "abs"
        elif kind in ("wall", "symmetry"):
            # BoundCondWall at first order for both; at second order
            # BoundCondSymmetryO2 is BoundCondWallO2 in all but the name
            bcs[num - 1] = "wall"
        elif kind == "outflow":  # BoundCondOutflow.cpp:41-52: <dataOutflow p0="..."/>
            bcs[num - 1] = "outflow"
            p0s[num - 1] = _f(_child(b, "dataOutflow", "boundCond outflow"), "p0", "boundCond outflow")
            bc_tr[num - 1] = _named_values(b.find("dataOutflow"), trs, "boundCond outflow", True)
        elif kind == "injection":
            bcs[num - 1] = "injection"
            injs[num - 1] = _injection(b, case["eos"])
            if trs:
                bc_tr[num - 1] = _named_values(b.find("dataInjection"), trs, "boundCond injection", True)
        elif kind == "tank":
            bcs[num - 1] = "tank"
            tanks[num - 1] = _tank(b, case["eos"])
            if trs:
                bc_tr[num - 1] = _named_values(b.find("dataTank"), trs, "boundCond tank", True)
        else:
            raise DeckError(f"boundCond type {kind!r} is not read")
    if model == "eulerhomogeneous":
        # ModEulerHomogeneous has only the interior Riemann solver (Model.h:
        # 93-119 end the run for walls, outflows, injections and tanks), and its
        # MUSCL form starts the first predictor's Newton solve from a pressure
        # PhaseEulerHomogeneous's reader leaves unset (its results are NaN)
        if any(b != "abs" for b in bcs):
            raise DeckError("EulerHomogeneous takes transmissive boundaries only")
        if case["order"] != 1:
            raise DeckError("EulerHomogeneous runs at first order (the reference's MUSCL form reads an unset pressure)")
    case["bc"] = bcs
    case["bc_p0"] = p0s
    if any(injs):
        case["bc_inj"] = injs
    if any(tanks):
        case["bc_tank"] = tanks
    if any(v is not None for v in bc_tr):
        case["bc_tr"] = bc_tr

    X, Y, Z = _centres(case["cells"], case["lengths"])
    Xg, Yg, Zg = X[None, None, :], Y[None, :, None], Z[:, None, None]
    nf = 3 * nph + 3 if model != "euler" else 6
    if trs:
        nf += 1 + len(trs)  # E (not read), then the scalars
    if model in ("thermaleq", "eulerhomogeneous"):
        nf += 2  # E (not read), then the temperature
    grid = np.zeros((nf, case["cells"][2], case["cells"][1], case["cells"][0]))
    filled = np.zeros(grid.shape[1:], dtype=bool)
    pdoms = ci.find("physicalDomains")
    for k, dom in enumerate(pdoms.findall("domain") if pdoms is not None else []):
        where = f"domain {dom.get('name', k)}"
        sname = dom.get("state")
        if sname not in states:
            raise DeckError(f"{where}: unknown state {sname!r}")
        vals = _state_values(states[sname], model, nph, where, case["eos"], (case.get("liq", 0), case.get("vap", 1)))
        if trs:
            vals = vals + [0.0] + _named_values(states[sname], trs, where, False)
        m = _domain_mask(dom, Xg, Yg, Zg, where)
        for f, v in enumerate(vals):
            grid[f][m] = v
        filled |= m
    if not filled.all():
        raise DeckError("initialConditionsV4.xml: the domains leave cells without a state")
    return case, grid.reshape(nf, -1)


def _injection(b, eos):
    """BoundCondInj's reader (BoundCondInj.cpp:46-150): <dataInjection m0>, then
    one <dataFluid density pressure> for one phase, or per phase a <dataFluid
    EOS alpha density pressure> under <fluidsProp>, the fractions in [0, 1]
    and summing to 1 within 1e-6 (they are divided by their sum when lowered)."""
    where = "boundCond injection"
    m0 = _f(_child(b, "dataInjection", where), "m0", where)
    nph = len(eos)
    if nph == 1:
        fl = _child(b, "dataFluid", where)
        return {"m0": m0, "alpha": [1.0], "rho": [_f(fl, "density", where)], "p": [_f(fl, "pressure", where)]}
    alpha, rho, p = [None] * nph, [None] * nph, [None] * nph
    fluids = _child(b, "fluidsProp", where).findall("dataFluid")
    names = [e.get("file") or f"{e['type']}_{e['name']}.xml" for e in eos]
    for fl in fluids:
        name = fl.get("EOS")
        if name not in names:
            raise DeckError(f"{where}: unknown EOS {name!r}")
        k = names.index(name)
        alpha[k], rho[k], p[k] = _f(fl, "alpha", where), _f(fl, "density", where), _f(fl, "pressure", where)
    if len(fluids) != nph or None in alpha:
        raise DeckError(f"{where}: {len(fluids)} fluids for {nph} phases")
    if any(a < 0.0 or a > 1.0 for a in alpha):
        raise DeckError(f"{where}: alpha should be in [0,1]")
    if abs(sum(alpha) - 1.0) > 1e-6:
        raise DeckError(f"{where}: sum of alpha should be 1")
    return {"m0": m0, "alpha": alpha, "rho": rho, "p": p}


def _tank(b, eos):
    """BoundCondTank's reader (BoundCondTank.cpp:46-144): <dataTank p0 T0>,
    then for several phases one <dataFluid EOS alpha|massFrac> per phase under
    <fluidsProp>, all alpha or all massFrac, each in [0, 1] and summing to 1
    within 1e-6 (divided by their sum when lowered)."""
    where = "boundCond tank"
    dt = _child(b, "dataTank", where)
    tk = {"p0": _f(dt, "p0", where), "T0": _f(dt, "T0", where)}
    nph = len(eos)
    if nph == 1:
        return tk
    names = [e.get("file") or f"{e['type']}_{e['name']}.xml" for e in eos]
    fluids = _child(b, "fluidsProp", where).findall("dataFluid")
    if len(fluids) != nph:
        raise DeckError(f"{where}: {len(fluids)} fluids for {nph} phases")
    kinds = {("massFrac" if fl.get("massFrac") is not None else "alpha") for fl in fluids}
    for fl in fluids:
        if (fl.get("alpha") is None) == (fl.get("massFrac") is None):
            raise DeckError(f"{where}: one of alpha, massFrac is required")
    if len(kinds) != 1:
        raise DeckError(f"{where}: only one of alpha, massFrac is required")
    key = kinds.pop()
    frac = [None] * nph
    for fl in fluids:
        name = fl.get("EOS")
        if name not in names:
            raise DeckError(f"{where}: unknown EOS {name!r}")
        frac[names.index(name)] = _f(fl, key, where)
    if None in frac:
        raise DeckError(f"{where}: {nph} phases need one fluid each")
    if any(v < 0.0 or v > 1.0 for v in frac):
        raise DeckError(f"{where}: {key} should be in [0,1]")
    if abs(sum(frac) - 1.0) > 1e-6:
        raise DeckError(f"{where}: sum of {key} should be 1")
    tk[key] = frac
    return tk


def read_vtr(path):
    """{name: values} of the CellData arrays of a result file in the
    reference's VTK XML format: ASCII as printed, or base64 (a 32-bit byte
    count, then little-endian Float32 data, each block encoded on its own,
    Output::ecritJeuDonnees, Output.cpp:219-266) widened to double.  Vector
    arrays come as (ncell, components)."""
    import base64
    import xml.etree.ElementTree as ET
    root = ET.parse(path).getroot()
    cd = root.find(".//CellData")
    if cd is None:
        raise DeckError(f"{Path(path).name}: no <CellData>")
    out = {}
    for da in cd.findall("DataArray"):
        text = (da.text or "").strip()
        if da.get("format") == "binary":
            head = base64.b64decode(text[:8])
            nbytes = int(np.frombuffer(head, dtype="<i4")[0])
            v = np.frombuffer(base64.b64decode(text[8:]), dtype="<f4")[: nbytes // 4].astype(np.float64)
        else:
            v = np.array(text.split(), dtype=np.float64)
        nc = int(da.get("NumberOfComponents", "1"))
        out[da.get("Name")] = v.reshape(-1, nc) if nc > 1 else v
    return out


def _array(arrs, prefix, where):
    keys = [k for k in arrs if k.startswith(prefix)]
    if not keys:
        raise DeckError(f"{where}: no array {prefix}*")
    return arrs[keys[0]]


def resume_point(case, rundir):
    """Run::resumeSimulation (src/Run.cpp:164-190) for result file N =
    case["resume"] under RUNDIR/results/<run>/: Output::readInfos
    (Output.cpp:370-405) finds row N of infoCalcul.out (one CPU) and takes its
    iteration, physical time and dt; OutputXML::readResults (OutputXML.cpp:
    94-200) reads the file's phase and mixture arrays; the primitives follow
    Phase*::setScalar / Mix*::setScalar, with every Kapila phase at the
    mixture pressure (ModKapila::fulfillState, ModKapila.cpp:87-99).
    Returns (host state, iteration, t, dt, infoCalcul.out lines kept)."""
    n = case["resume"]
    if case["output"]["format"] != "XML":
        raise DeckError("readResutls not available for requested output format")  # Output.h:79
    if case["output"]["binary"]:
        raise DeckError("resuming on binary results file not available")  # Output::getJeuDonnees, Output.cpp:279
    rdir = Path(rundir) / "results" / case["name"]
    info = rdir / "infoCalcul.out"
    if not info.exists():
        raise DeckError(f"resume: missing {info}")
    lines = info.read_text().split("\n")
    if not lines or lines[0].split()[:1] != ["1"]:
        raise DeckError("resume simulation not possible - number of CPU differs from read files")
    row = None
    for k, ln in enumerate(lines[1:], start=1):
        f = ln.split()
        if f and int(f[0]) == n:
            row, kept = f, lines[: k + 1]
            break
    if row is None:
        raise DeckError("resume simulation not possible - check file 'infoCalcul.out'")
    it, t, dt = int(row[1]), float(row[2]), float(row[4])
    where = f"result_CPU0_AMR0_TIME{n}.vtr"
    arrs = read_vtr(rdir / where)
    model, nph = case["model"], len(case["eos"])
    ncell = int(np.prod(case["cells"]))
    ntr = len(case.get("transports", []))
    te = model == "thermaleq"
    state = np.zeros((3 * nph + 4 + ntr + te if (ntr or te) else 3 * nph + 3, ncell))
    # the scalars T1, T2, ... (OutputXML::readResults, OutputXML.cpp:196-203)
    for t in range(ntr):
        if f"T{t + 1}" not in arrs:
            raise DeckError(f"{where}: no array T{t + 1}")
        state[3 * nph + 4 + t] = arrs[f"T{t + 1}"]
    if model == "euler":
        state[0] = 1.0
        state[1] = _array(arrs, "F0_Density_", where)
        state[2] = _array(arrs, "F0_Pressure_", where)
        vel = _array(arrs, "F0_Velocity_", where)
    elif te:
        # PhaseThermalEq::setScalar (alpha, density), MixThermalEq::setScalar
        # (density, pressure, temperature) and setVector (velocity)
        pm = _array(arrs, "Pressure_Mixture", where)
        for k in range(nph):
            state[3 * k] = _array(arrs, f"F{k}_Alpha_", where)
            state[3 * k + 1] = _array(arrs, f"F{k}_Density_", where)
            state[3 * k + 2] = pm
        state[3 * nph + 4] = _array(arrs, "Temperature_Mixture", where)
        vel = _array(arrs, "Velocity_Mixture", where)
    else:
        pm = _array(arrs, "Pressure_Mixture", where)
        for k in range(nph):
            state[3 * k] = _array(arrs, f"F{k}_Alpha_", where)
            state[3 * k + 1] = _array(arrs, f"F{k}_Density_", where)
            state[3 * k + 2] = pm if model == "kapila" else _array(arrs, f"F{k}_Pressure_", where)
        vel = _array(arrs, "Velocity_Mixture", where)
    state[3 * nph:3 * nph + 3] = np.asarray(vel).reshape(ncell, 3).T
    return state, it, t, dt, kept


def _run_with_output(ctx, case, outdir, resume=None):
    """March to the deck's end, writing a result file at t = 0 and after every
    iterFreq steps (iteration control) or at the first step reaching each
    multiple of timeFreq (physical-time control), as the reference does, plus
    collection.pvd and infoCalcul.out (Output::saveInfos, Output.cpp:337-364:
    the CPU count, then per result file its number, the iteration, the
    physical time, the computation time in ms and the dt of the last step,
    1e-15 at t = 0: Run::solver writes before it updates dt, Run.cpp:234-270)."""
    import time
    from . import output
    o = case["output"]
    outdir.mkdir(parents=True, exist_ok=True)
    prec = o["precision"] or 6  # the stream's default precision when the deck sets none
    t0 = time.perf_counter()
    info = open(outdir / "infoCalcul.out", "w")
    if resume is None:
        files, it0, ct0 = [], 0, 0
        info.write("1\n")
    else:
        # the resumed run keeps infoCalcul.out up to row N (Output::readInfos)
        # and numbers its files from N + 1
        n0, it0, kept, _ = resume
        files = [outdir / f"result_CPU0_AMR0_TIME{k}.vtr" for k in range(n0 + 1)]
        info.write("\n".join(kept) + "\n")
        # the computation time continues from row N's (timeStats::setCompTime, Output.cpp:408)
        ct0 = int(kept[-1].split()[3])

    gnu = o["format"] == "GNU"
    cuts = [dict(cu, n=0) for cu in case.get("cuts", [])]
    if cuts:
        (outdir / "cuts").mkdir(exist_ok=True)

    def dump(st=None):
        if gnu:  # OutputGNU::ecritSolution: the file, then the script over all files so far
            f = outdir / f"result_CPU0_TIME{len(files)}.out"
            output.write_gnu(f, case, ctx.get_state())
            output.write_gnu_script(outdir / "visualisation.gnu", case, len(files) + 1)
        else:
            f = outdir / f"result_CPU0_AMR0_TIME{len(files)}.vtr"
            output.write_vtr(f, case, ctx.get_state(), binary=o["binary"], precision=prec)
        # cuts (Run.cpp:152, 259): after each result file, one file per cut
        # numbered by the cut's own count from 0 (a resumed run too); cuts of
        # one type share the file name, so the last of a type is what remains
        for cu in cuts:
            state = ctx.get_state()
            output.write_cut(outdir / "cuts" / f"{cu['type']}_CPU0_TIME{cu['n']}.out", case, state, cu)
            output.write_gnu_script(outdir / "cuts" / f"visualisation{cu['type']}.gnu", case, cu["n"] + 1,
                                    dim=1 if cu["type"] == "cut1D" else 2, stem=cu["type"])
            cu["n"] += 1
        it, t, dt = (0, 0.0, 1.e-15) if st is None else (it0 + st.steps, st.t, st.dt_last)
        ms = ct0 + int((time.perf_counter() - t0) * 1e3) if st is not None else 0
        info.write(f"{len(files)} {it} {t:.{prec}g} {ms} {dt:.{prec}g}\n")
        files.append(f)
    # probes (OutputProbeGNU): each file emptied at the start with its script;
    # a row at t = 0 (not on resume), then after every step that reaches the
    # next acquisition time, which then advances by acqFreq (Run.cpp:264-266)
    probes = []
    if case.get("probes"):
        (outdir / "probes").mkdir(exist_ok=True)
    for pr in case.get("probes", []):
        pf = outdir / "probes" / f"{pr['name']}.out"
        pf.write_text("")
        output.write_gnu_script(outdir / "probes" / f"visualisation{pr['name']}.gnu", case, 1, dim=0, stem=pr["name"])
        probes.append({"file": pf, "cell": output.probe_cell(case, pr["vertex"]), "freq": pr["acq_freq"], "next": 0.0})

    def acquire(t, state, first=False):
        for pb in probes:
            if first or pb["next"] <= t:
                with open(pb["file"], "a") as fh:
                    fh.write(output.probe_row(case, state, t, pb["cell"]))
                pb["next"] += pb["freq"]

    def advance(nsteps=0, t_final=None):
        """ctx.run for nsteps steps or up to t_final, one step at a time while
        probes are to be sampled."""
        if not probes:
            return ctx.run(nsteps=nsteps) if t_final is None else ctx.run(t_final=t_final)
        while True:
            stp = ctx.run(nsteps=1) if t_final is None else ctx.run(t_final=t_final, max_steps=1)
            acquire(stp.t, ctx.get_state())
            if t_final is None:
                nsteps -= 1
                if nsteps <= 0:
                    return stp
            elif stp.t >= float(np.float32(t_final)):
                return stp

    st = None
    if resume is None:
        dump()
        if probes:
            acquire(0.0, ctx.get_state(), first=True)
    if case["time_mode"] == "iterations":
        # a result file whenever the iteration is a multiple of iterFreq
        # (Run.cpp:239-241), none at a last iteration that is not; the loop
        # takes at least one step, as Run::solver does, also for a run
        # resumed at or beyond its step count
        step = o["iter_freq"] if o["iter_freq"] > 0 else case["steps"]
        done = it0
        while True:
            n = max(1, min(step - done % step, case["steps"] - done))
            st = advance(nsteps=n)
            done = it0 + st.steps
            if done % step == 0:
                dump(st)
            if done >= case["steps"]:
                break
    else:
        # Run::solver (Run.cpp:208, 241-243): a result file at each step with
        # t > printSuivante (double), which starts at the run's first time plus
        # timeFreq and grows by timeFreq at each file; both times are floats
        # (Input.cpp:164-166).  The run ends at the first t >= totalTime.
        tfin = float(np.float32(case["t_final"]))
        ff = float(np.float32(o["time_freq"]))
        print_next = (resume[3] if resume is not None else 0.0) + ff
        t = resume[3] if resume is not None else 0.0
        first = True
        while first or t < tfin:  # at least one step (Run::solver)
            first = False
            # the device stops at the first t >= a float target: the largest
            # float not above print_next, then single steps to t > print_next
            low = np.float32(print_next)
            if float(low) > print_next:
                low = np.nextafter(low, np.float32(-np.inf))
            target = min(float(low), tfin)
            st = advance(t_final=target) if t < target else advance(nsteps=1)
            while not (st.t > print_next) and st.t < tfin:
                st = advance(nsteps=1)
            t = st.t
            if t > print_next:
                dump(st)
                print_next += ff
        if st is None:
            st = ctx.sync()
    if not gnu:
        output.write_collection(outdir / "collection.pvd", files)
    info.close()
    return st


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("rundir")
    ap.add_argument("--out", default="final.npz")
    ap.add_argument("--precision", choices=["fp64", "fp32s"], default="fp64")
    ap.add_argument("--vtk", action="store_true",
                    help="write VTK result files at the deck's output frequency under RUNDIR/results/<run>/")
    args = ap.parse_args(argv)
    from .capi import Context
    case, state = load(args.rundir)
    case["precision"] = args.precision
    it0, resume = 0, None
    if case.get("resume"):
        state, it0, t0, dt0, kept = resume_point(case, args.rundir)
        resume = (case["resume"], it0, kept, t0)
    with Context(case) as ctx:
        ctx.set_state(state)
        if resume is not None:
            ctx.set_clock(t0, dt0)
        if args.vtk:
            st = _run_with_output(ctx, case, Path(args.rundir) / "results" / case["name"], resume)
        elif case["time_mode"] == "iterations":
            # Run::solver takes at least one step
            st = ctx.run(nsteps=max(1, case["steps"] - it0))
        else:
            st = ctx.run(t_final=case["t_final"])
        out = ctx.get_state()
    steps = it0 + st.steps
    np.savez(args.out, state=out, steps=steps, t=st.t, case_json=json.dumps(case))
    print(f"{case['name']}: iteration {steps}, t = {st.t!r} -> {args.out}")


if __name__ == "__main__":
    main()
