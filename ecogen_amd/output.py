"""Results in the reference's VTK XML format, so that device results load in
the same post-processing as the reference's own.

``write_vtr`` writes one RectilinearGrid file of a Cartesian state as
``OutputXML::ecritSolutionXML`` does (``src/InputOutput/OutputXML.cpp:496-560``
for the mesh, ``:353-480`` for the cell data): node coordinates, then per phase
the scalars of ``Phase*::returnNameScalar`` named ``F<k>_<name>_<EOS file>``,
the mixture scalars and velocity (multiphase models), and ``gradRho``
(``Cell::getGradient``, ``src/Cell.cpp:525-682, 880-887``).  Arrays are typed
Float32 as in the reference; ASCII output prints the double values with the
deck's precision, binary output is base64 (a 32-bit byte count, then the
little-endian Float32 data, ``Output::ecritJeuDonnees``, ``Output.cpp:219-266``).
``write_collection`` writes the ``collection.pvd`` time series.

Every array is computed from the stored state.  With MultiP pressure
relaxation the reference prints instead the phase temperatures and the
mixture pressure it cached before its last relaxation (the relaxation does
not refresh them); there the values differ by what the relaxation moved.
"""
from __future__ import annotations

import base64
import struct
from pathlib import Path

import numpy as np


def _eos_file(e):
    return e.get("file") or f"{e['type']}_{e['name']}.xml"


def _temperature(e, rho, p, eps):
    """Eos{IG,SG}::computeTemperature (EosIG.cpp:90-93, EosSG.cpp:93-96)."""
    r = np.maximum(rho, eps)
    if e["type"] == "SG":
        return (p + e["pinf"]) / (e["gamma"] - 1.0) / r / e["cv"]
    return p / (e["gamma"] - 1.0) / r / e["cv"]


def _positions(case):
    """Cell centres (i + 0.5) dX and widths dX per axis (MeshCartesian)."""
    n, L = case["cells"], case["lengths"]
    d = [L[a] / n[a] for a in range(3)]
    return [(np.arange(n[a]) + 0.5) * d[a] for a in range(3)], d


def grad_rho(case, rho):
    """|grad rho| of Cell::getGradient on a Cartesian grid: per interior face
    (cd - cg)/|xR - xL| times the normal, weighted by the centre distance along
    each axis; every boundary face (any type, for the density) adds twice the
    centre-to-face distance to the weights only (outflow, injection and tank
    faces nothing)."""
    n = case["cells"]
    dim = sum(1 for c in n if c > 1)
    pos, dx = _positions(case)
    r = rho.reshape(n[2], n[1], n[0])
    g = [np.zeros_like(r) for _ in range(3)]
    w = [np.zeros_like(r) for _ in range(3)]
    axis_of = {0: 2, 1: 1, 2: 0}  # array axis of each direction in [k, j, i]
    for d in range(dim):
        ax = axis_of[d]
        p = pos[d]
        dist = np.abs(p[:-1] - p[1:])        # |xL - xR| of each interior face (Element::distance)
        shape = [1, 1, 1]
        shape[ax] = -1
        dist = dist.reshape(shape)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        lo, hi = tuple(lo), tuple(hi)
        face = ((r[hi] - r[lo]) / dist) * 1.0 * dist   # normal component 1, times |distance along d|
        # accumulation order of a cell's faces: the minus face, then the plus face
        gm = np.zeros_like(r)
        gp = np.zeros_like(r)
        gp[lo] = face                      # the plus face of the left cell
        gm[hi] = face                      # the minus face of the right cell
        g[d] = (g[d] + gm) + gp
        wm = np.zeros_like(r)
        wp = np.zeros_like(r)
        wp[lo] = np.broadcast_to(dist, face.shape)
        wm[hi] = np.broadcast_to(dist, face.shape)
        w[d] = (w[d] + wm) + wp
    # boundary faces after the interior ones (X-, X+, Y-, ...): twice the
    # centre-to-face distance, the X+ face lying at x_{n-1} + dX/2; outflow,
    # injection and tank faces (whoAmI 3, 4, 5) have no branch there and add nothing
    bcs = case.get("bc", ["abs"] * 6)
    silent = ("outflow", "injection", "tank")
    for d in range(dim):
        ax = axis_of[d]
        p = pos[d]
        first = [slice(None)] * 3
        last = [slice(None)] * 3
        first[ax] = slice(0, 1)
        last[ax] = slice(-1, None)
        if bcs[2 * d] not in silent:
            w[d][tuple(first)] += np.abs(p[0] - 0.0) * 2.0
        if bcs[2 * d + 1] not in silent:
            w[d][tuple(last)] += np.abs(p[-1] - (p[-1] + 0.5 * dx[d])) * 2.0
    comp = []
    for d in range(3):
        ww = np.where(w[d] <= 1e-12, 1.0, w[d])
        comp.append(g[d] / ww)
    return np.sqrt(comp[0] * comp[0] + comp[1] * comp[1] + comp[2] * comp[2]).reshape(-1)


def cell_data(case, state):
    """[(name, components, values)] of the reference's output for ``state``
    (ecogen.h host layout: per phase alpha, rho, p; u, v, w; E optional)."""
    model = case["model"]
    eos = case["eos"]
    eps = 1e-15 if (model == "kapila" and case.get("alpha_null")) else 0.0
    out = []
    if model == "euler":
        e = eos[0]
        rho, p = state[1], state[2]
        vel = np.stack([state[3], state[4], state[5]], axis=1)
        f = _eos_file(e)
        out += [(f"F0_Density_{f}", 1, rho), (f"F0_Pressure_{f}", 1, p),
                (f"F0_Temperature_{f}", 1, _temperature(e, rho, p, eps)), (f"F0_Velocity_{f}", 3, vel.reshape(-1))]
        out.append(("gradRho", 1, grad_rho(case, rho)))
        return out
    nph = len(eos)
    if model == "eulerhomogeneous":
        # PhaseEulerHomogeneous::returnNameScalar (Alpha, Density, Mass
        # Fraction), MixEulerHomogeneous::returnNameScalar / returnNameVector
        # (PhaseEulerHomogeneous.cpp:66-96, MixEulerHomogeneous.cpp:247-290)
        rho_m = np.zeros(state.shape[1])
        for k in range(nph):
            rho_m = rho_m + state[3 * k] * state[3 * k + 1]
        for k, e in enumerate(eos):
            f = _eos_file(e)
            out += [(f"F{k}_Alpha_{f}", 1, state[3 * k]), (f"F{k}_Density_{f}", 1, state[3 * k + 1]),
                    (f"F{k}_Mass Fraction_{f}", 1, state[3 * k] * state[3 * k + 1] / rho_m)]
        vel = np.stack([state[3 * nph], state[3 * nph + 1], state[3 * nph + 2]], axis=1)
        out += [("Density_Mixture", 1, rho_m), ("Pressure_Mixture", 1, state[2]),
                ("Temperature_Mixture", 1, state[3 * nph + 4]), ("Velocity_Mixture", 3, vel.reshape(-1))]
        out.append(("gradRho", 1, grad_rho(case, rho_m)))
        return out
    if model == "thermaleq":
        # PhaseThermalEq::returnNameScalar (Alpha, Density), MixThermalEq::
        # returnNameScalar / returnNameVector (PhaseThermalEq.cpp:82-104,
        # MixThermalEq.cpp:357-403)
        rho_m = np.zeros(state.shape[1])
        for k in range(nph):
            rho_m = rho_m + state[3 * k] * state[3 * k + 1]
        for k, e in enumerate(eos):
            f = _eos_file(e)
            out += [(f"F{k}_Alpha_{f}", 1, state[3 * k]), (f"F{k}_Density_{f}", 1, state[3 * k + 1])]
        vel = np.stack([state[3 * nph], state[3 * nph + 1], state[3 * nph + 2]], axis=1)
        out += [("Density_Mixture", 1, rho_m), ("Pressure_Mixture", 1, state[2]),
                ("Temperature_Mixture", 1, state[3 * nph + 4]), ("Velocity_Mixture_norm", 3, vel.reshape(-1))]
        out.append(("gradRho", 1, grad_rho(case, rho_m)))
        return out
    rho_m = np.zeros(state.shape[1])
    p_m = np.zeros(state.shape[1])
    for k in range(nph):
        rho_m = rho_m + state[3 * k] * state[3 * k + 1]
        p_m = p_m + state[3 * k] * state[3 * k + 2]
    for k, e in enumerate(eos):
        a, r, p = state[3 * k], state[3 * k + 1], state[3 * k + 2]
        f = _eos_file(e)
        T = _temperature(e, r, p, eps)
        if model == "kapila":
            Y = a * r / np.maximum(rho_m, eps)
            # PhaseKapila::returnScalar (PhaseKapila.cpp:115-138) clamps for output
            out += [(f"F{k}_Alpha_{f}", 1, np.where(a < 1e-20, 0.0, a)),
                    (f"F{k}_Density_{f}", 1, np.where(r < 1e-20, 1e-20, r)),
                    (f"F{k}_Temperature_{f}", 1, np.where(T < 1e-20, 1e-20, T)),
                    (f"F{k}_Mass fraction_{f}", 1, np.where(Y < 1e-20, 0.0, Y))]
        else:
            Y = a * r / rho_m
            out += [(f"F{k}_Alpha_{f}", 1, a), (f"F{k}_Density_{f}", 1, r), (f"F{k}_Temperature_{f}", 1, T),
                    (f"F{k}_Mass fraction_{f}", 1, Y), (f"F{k}_Pressure_{f}", 1, p)]
    vel = np.stack([state[3 * nph], state[3 * nph + 1], state[3 * nph + 2]], axis=1)
    out += [("Density_Mixture", 1, rho_m), ("Pressure_Mixture", 1, p_m), ("Velocity_Mixture", 3, vel.reshape(-1))]
    # transported scalars T1, T2, ... (OutputXML.cpp:435-447), stored after E
    for t in range(len(case.get("transports", []))):
        out.append((f"T{t + 1}", 1, state[3 * nph + 4 + t]))
    out.append(("gradRho", 1, grad_rho(case, rho_m)))
    return out


def _ascii(values, precision):
    return "".join(f"{v:.{precision}g} " for v in np.asarray(values, dtype=np.float64))


def _binary(values):
    data = np.asarray(values, dtype="<f4").tobytes()
    return (base64.b64encode(struct.pack("<i", len(data))) + base64.b64encode(data)).decode()


def write_vtr(path, case, state, binary=False, precision=17):
    """One result file of ``state`` (host layout) in the reference's format."""
    n = case["cells"]
    pos, dx = _positions(case)
    fmt = "binary" if binary else "ascii"

    def enc(v):
        return _binary(v) if binary else _ascii(v, precision)
    lines = ['<?xml version="1.0"?>',
             '<VTKFile type="RectilinearGrid" version="0.1" byte_order="LittleEndian">',
             f'  <RectilinearGrid WholeExtent="0 {n[0]} 0 {n[1]} 0 {n[2]}">',
             f'    <Piece Extent="0 {n[0]} 0 {n[1]} 0 {n[2]}">',
             "      <Coordinates>"]
    for a in range(3):
        # MeshCartesian::recupereCoord: x_i - dX/2 for every cell, then x_{n-1} + dX/2
        nodes = np.append(pos[a] - 0.5 * dx[a], pos[a][-1] + 0.5 * dx[a])
        lead = "\n          " if not binary else "\n"
        lines.append(f'        <DataArray type="Float32" format="{fmt}">{lead}{enc(nodes)}')
        lines.append("        </DataArray>")
    lines += ["      </Coordinates>", "      <CellData>"]
    for name, nc, vals in cell_data(case, np.asarray(state, dtype=np.float64)):
        comp = f' NumberOfComponents="{nc}"' if nc > 1 else ""
        lines.append(f'        <DataArray type="Float32" Name="{name}"{comp} format="{fmt}">\n{enc(vals)}')
        lines.append("        </DataArray>")
    lines += ["      </CellData>", "    </Piece>", "  </RectilinearGrid>", "</VTKFile>", ""]
    Path(path).write_text("\n".join(lines))


def write_collection(path, files):
    """collection.pvd: one DataSet per written file, timestep = its index
    (OutputXML::ecritCollectionXML, OutputXML.cpp:323-350)."""
    rows = "".join(f'        <DataSet timestep="{t}" part="0" file="{Path(f).name}"/>\n' for t, f in enumerate(files))
    Path(path).write_text('<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1" byte_order="LittleEndian" '
                          'compressor="vtkZLibDataCompressor">\n    <Collection>\n' + rows +
                          "    </Collection>\n</VTKFile>\n")


def _gnu_columns(case, state):
    """(title, values) per column after the coordinates: the phase and mixture
    variables of cell_data, vectors as their norm sqrt(x^2 + y^2 + z^2)
    (Phase::printPhase, Mixture::printMixture, Coord::norm), then the AMR
    level and the refinement indicator xi, both 0 on a fixed mesh
    (Cell::printGnuplotAMR, src/Cell.cpp:1523-1556)."""
    cols = []
    for name, nc, vals in cell_data(case, np.asarray(state, dtype=np.float64)):
        if name == "gradRho":
            continue
        title = name.split("_", 1)[1] if name[:1] == "F" and name[1:2].isdigit() else name
        if name[:1] == "T" and name[1:].isdigit():
            title = f"Transport{name[1:]}"  # OutputGNU.cpp:140-145
        if nc > 1:
            v = np.asarray(vals).reshape(-1, nc)
            vals = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2])
        cols.append((title, vals))
    n = int(np.prod(case["cells"]))
    cols += [("Niveau AMR", np.zeros(n)), ("Xi", np.zeros(n))]
    return cols


def write_gnu(path, case, state):
    """One result file of OutputGNU::ecritSolution (OutputGNU.cpp:59-75): per
    cell in cell order its centre (x, then y and z by dimension) and the
    columns of _gnu_columns, each value at the stream's default 6 significant
    digits followed by a space, one line per cell, then an empty line."""
    n = case["cells"]
    dim = sum(1 for c in n if c > 1)
    pos, _ = _positions(case)
    X = np.broadcast_to(pos[0][None, None, :], (n[2], n[1], n[0])).reshape(-1)
    Y = np.broadcast_to(pos[1][None, :, None], (n[2], n[1], n[0])).reshape(-1)
    Z = np.broadcast_to(pos[2][:, None, None], (n[2], n[1], n[0])).reshape(-1)
    data = [X, Y, Z][:dim] + [v for _, v in _gnu_columns(case, state)]
    rows = np.stack(data, axis=1)
    text = "".join("".join(f"{v:.6g} " for v in r) + "\n" for r in rows)
    Path(path).write_text(text + "\n")


def write_gnu_script(path, case, nfiles, dim=None, stem="result"):
    """visualisation.gnu of OutputGNU::ecritScriptGnuplot (OutputGNU.cpp:79-165)
    for result files 0 .. nfiles-1 of one CPU (<stem>_CPU0_TIME<t>.out): one
    plot (1-D) or splot with contours (2-D) per column; ``dim`` 0 is a probe's
    time history (<stem>.out, against t).  The reference writes none in 3-D."""
    if dim is None:
        dim = sum(1 for c in case["cells"] if c > 1)
    if dim == 3:
        raise ValueError("no gnuplot script in 3-D (OutputGNU.cpp:110)")
    out = ["reset", "set style data lines", "set nokey", ""]
    out.append("set xlabel 't (s)'" if dim == 0 else "set xlabel 'x (m)'")
    if dim == 2:
        out += ["set surface", "set dgrid3d 50,50", "set contour base", "set cntrparam levels 25", "show contour"]
    index = 2 if dim == 0 else 1 + dim
    nf = 3 * len(case["eos"]) + 4 + len(case.get("transports", [])) + (case["model"] in ("thermaleq", "eulerhomogeneous"))
    for title, _ in _gnu_columns(case, np.ones((nf, int(np.prod(case["cells"]))))):
        out.append(f"set title '{title}'")
        if dim == 0:
            out.append(f'plot "{stem}.out" u 1:{index}')
        else:
            use = f"u 1:{index}" if dim <= 1 else f"u 1:2:{index}"
            files = [f'"{stem}_CPU0_TIME{t}.out" {use}' for t in range(nfiles)]
            out.append(("plot " if dim <= 1 else "splot ") + ",\\\n ".join(files))
        out += ["pause(-1)", ""]
        index += 1
    Path(path).write_text("\n".join(out) + "\n")


def _centres(case):
    n = case["cells"]
    pos, _ = _positions(case)
    X = np.broadcast_to(pos[0][None, None, :], (n[2], n[1], n[0])).reshape(-1)
    Y = np.broadcast_to(pos[1][None, :, None], (n[2], n[1], n[0])).reshape(-1)
    Z = np.broadcast_to(pos[2][:, None, None], (n[2], n[1], n[0])).reshape(-1)
    return X, Y, Z


def _lcfl(case):
    """ElementCartesian's lCFL on a uniform grid (MeshCartesian.cpp:256-262):
    the smallest cell width over the non-degenerate directions, times 0.6 in
    2-D and 3-D."""
    n, L = case["cells"], case["lengths"]
    l = 1.e10
    for a in range(3):
        if n[a] != 1:
            l = min(l, L[a] / n[a])
    return l * 0.6 if sum(1 for c in n if c > 1) > 1 else l


def _rows(coords, cols, idx):
    data = list(coords) + [np.asarray(v)[idx] for _, v in cols]
    rows = np.stack(data, axis=1) if data else np.zeros((len(idx), 0))
    return "".join("".join(f"{v:.6g} " for v in r) + "\n" for r in rows)


def _unit(v):
    """Coord::normalized: each component divided by the norm."""
    n = np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    return [v[0] / n, v[1] / n, v[2] / n]


def _plane_base(point, normal):
    """GOPlan::createBase (GOPlan.cpp:82-106): a tangent from the point to the
    plane's intersection with the last axis whose normal component is at
    least 1e-6 (z, then y, then x), normalized; binormal = normal x tangent."""
    M, n = point, normal
    if abs(n[2]) >= 1e-6:
        N = [0.0, 0.0, (M[0] * n[0] + M[1] * n[1]) / n[2] + M[2]]
    elif abs(n[1]) >= 1e-6:
        N = [0.0, (M[0] * n[0] + M[2] * n[2]) / n[1] + M[1], 0.0]
    else:
        N = [(M[1] * n[1] + M[2] * n[2]) / n[0] + M[0], 0.0, 0.0]
    t = _unit([N[0] - M[0], N[1] - M[1], N[2] - M[2]])
    b = [n[1] * t[2] - n[2] * t[1], n[2] * t[0] - n[0] * t[2], n[0] * t[1] - n[1] * t[0]]
    return t, b


def cut_cells(case, cut):
    """The cells a cut prints, in cell order, and their coordinates on it
    (Cell::printGnuplotAMR with Element::traverseObjet, src/Cell.cpp:1523-1556,
    src/Meshes/Element.cpp:182-186): a cell is printed when the object's
    distance to its centre is below lCFL.

    * cut1D, a line through ``vertex`` along the normalized ``vec``
      (GOLine::distancePoint, GOLine.cpp:61-66): with w = centre - vertex,
      the distance is 0 where abs(w . dir) < 1e-6, else |w x dir|.  That
      source file calls the C library's int abs(int) (it has no using
      namespace std), so w . dir is truncated to an integer first: every
      cell less than 1 (in length units) along the line from the vertex,
      measured parallel to it, has distance 0 and is printed.  Its coordinate
      is dir . w (GOLine::projectionPoint).
    * cut2D, a plane through ``vertex`` with the normalized normal ``vec``
      (GOPlan::distancePoint, GOPlan.cpp:64-68): |w . n| with the FP64 abs;
      its coordinates are (tangent . w, binormal . w) (GOPlan::projectionPoint)."""
    X, Y, Z = _centres(case)
    P, v = cut["vertex"], _unit(cut["vec"])
    wx, wy, wz = X - P[0], Y - P[1], Z - P[2]
    lim = _lcfl(case)
    if cut["type"] == "cut1D":
        t = v[0] * wx + v[1] * wy + v[2] * wz
        cx, cy, cz = wy * v[2] - wz * v[1], wz * v[0] - wx * v[2], wx * v[1] - wy * v[0]
        dist = np.where(np.trunc(t) == 0.0, 0.0, np.sqrt(cx * cx + cy * cy + cz * cz))
        return np.nonzero(dist < lim)[0], [t]
    tg, bn = _plane_base(P, v)
    dist = np.abs(wx * v[0] + wy * v[1] + wz * v[2])
    return np.nonzero(dist < lim)[0], [tg[0] * wx + tg[1] * wy + tg[2] * wz, bn[0] * wx + bn[1] * wy + bn[2] * wz]


def write_cut(path, case, state, cut):
    """One file of OutputCutGNU::ecritSolution (OutputCutGNU.cpp:105-123):
    per printed cell its coordinates on the cut and the columns of the GNU
    result files, then an empty line."""
    idx, coords = cut_cells(case, cut)
    text = _rows([c[idx] for c in coords], _gnu_columns(case, np.asarray(state, dtype=np.float64)), idx)
    Path(path).write_text(text + "\n")


def probe_cell(case, vertex):
    """OutputProbeGNU::locateProbeInMesh (OutputProbeGNU.cpp:115-135): the
    first cell, in cell order, of least distance |centre - vertex|, the search
    stopping at the first within lCFL."""
    X, Y, Z = _centres(case)
    dx, dy, dz = X - vertex[0], Y - vertex[1], Z - vertex[2]
    dist = np.sqrt(dx * dx + dy * dy + dz * dz)
    lim = _lcfl(case)
    best, bi = 1.e12, 0
    for i, di in enumerate(dist):
        if di < best:
            best, bi = di, i
            if di <= lim:
                break
    return bi


def probe_row(case, state, t, cell):
    """One line of a probe file (OutputProbeGNU::ecritSolution,
    OutputProbeGNU.cpp:154-170): the physical time, then the cell's columns."""
    return f"{t:.6g} " + _rows([], _gnu_columns(case, state), np.array([cell]))
