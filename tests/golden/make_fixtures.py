"""Converts the reference's non-regression outputs into compact fixtures.

Reads NONREGRESSION/<case>/output of the MaHyCo tree (the reference's own test
data: EnSight Gold ASCII final states and 17-digit EnvSummation time histories,
NONREGRESSION/launch_test.sh:74-96) and writes tests/golden/<case>.npz holding:

  th_time                       time column (pins the whole dt sequence)
  th_<Var>_<env index>          EnvSummation value column, Var in CellMass,
                                Pressure, Density, InternalEnergy
  cell_<Var>                    final per-cell field in global cell order
                                (x fastest), printed with %12.5e by Arcane
  node_<Var>[_<c>]              final per-node field in global node order
  final_time                    ensight.case time value
  env_names                     environment names of the dataset

Parallel runs write one EnSight part per subdomain; parts are mapped back to
global cells through the part geometry (cell centroids, node coordinates).

Run inside a container that holds the reference tree:
    python tests/golden/make_fixtures.py /root/reference/NONREGRESSION
"""
import os
import re
import sys

import numpy as np

CASES = {
    # name: (dim, (nx, ny, nz), origin, length)
    "CAS_BiSodCaseX": (3, (100, 5, 5), (0, 0, 0), (1., .1, .1)),
    "CAS_BiSodCaseY": (3, (5, 100, 5), (0, 0, 0), (.1, 1., .1)),
    "CAS_BiSodCaseZ": (3, (5, 5, 100), (0, 0, 0), (.1, .1, 1.)),
    "CAS_SodCaseY": (3, (5, 100, 5), (0, 0, 0), (.1, 1., .1)),
    "CAS_SodCaseZ": (3, (5, 5, 100), (0, 0, 0), (.1, .1, 1.)),
    "CAS_Lag_BiSodCaseX": (3, (100, 5, 5), (0, 0, 0), (1., .1, .1)),
    "CAS_2D_Lag_SodCaseX": (2, (100, 5, 1), (0, 0, 0), (1., .1, 0.)),
    "CAS_2D_Sedov": (2, (60, 60, 1), (0, 0, 0), (1.2, 1.2, 0.)),
    "CAS_bitriplepoint_minmod": (2, (70, 30, 1), (0, 0, 0), (.07, .03, 0.)),
    "CAS_BiSodCaseX_ordre_3": (3, (100, 5, 5), (0, 0, 0), (1., .1, .1)),
    "CAS_BiSodCaseY_ordre_3": (3, (5, 100, 5), (0, 0, 0), (.1, 1., .1)),
    "CAS_BiSodCaseZ_ordre_3": (3, (5, 5, 100), (0, 0, 0), (.1, .1, 1.)),
    "CAS_SodCaseX_ordre_3": (3, (100, 5, 5), (0, 0, 0), (1., .1, .1)),
    "CAS_SodCaseY_ordre_3": (3, (5, 100, 5), (0, 0, 0), (.1, 1., .1)),
    "CAS_SodCaseZ_ordre_3": (3, (5, 5, 100), (0, 0, 0), (.1, .1, 1.)),
    "CAS_Lag_BiSodCaseY": (3, (5, 100, 5), (0, 0, 0), (.1, 1., .1)),
    "CAS_Lag_BiSodCaseZ": (3, (5, 5, 100), (0, 0, 0), (.1, .1, 1.)),
    "CAS_Lag_SodCaseX": (3, (100, 5, 5), (0, 0, 0), (1., .1, .1)),
    "CAS_Lag_SodCaseY": (3, (5, 100, 5), (0, 0, 0), (.1, 1., .1)),
    "CAS_Lag_SodCaseZ": (3, (5, 5, 100), (0, 0, 0), (.1, .1, 1.)),
    "CAS_2D_Lag_BiSodCaseX": (2, (100, 5, 1), (0, 0, 0), (1., .1, 0.)),
    "CAS_2D_Lag_BiSodCaseY": (2, (5, 100, 1), (0, 0, 0), (.1, 1., 0.)),
    "CAS_2D_Lag_SodCaseY": (2, (5, 100, 1), (0, 0, 0), (.1, 1., 0.)),
    "CAS_2D_Lag_SodCaseZ": (3, (5, 1, 100), (0, 0, 0), (.1, .1, 1.)),
    "CAS_Sedov": (3, (60, 60, 60), (0, 0, 0), (1.2, 1.2, 1.2)),
    "CAS_BiSodCaseX_pente_borne_minmod_debar_1": (3, (100, 5, 5), (0, 0, 0), (1., .1, .1)),
    "CAS_BiSodCaseX_pente_borne_minmod_debar_2": (3, (100, 5, 5), (0, 0, 0), (1., .1, .1)),
    "CAS_BiSodCaseX_pente_borne_superbee_debar_2": (3, (100, 5, 5), (0, 0, 0), (1., .1, .1)),
    "CAS_BiSodCaseY_pente_borne_minmod_debar_2": (3, (5, 100, 5), (0, 0, 0), (.1, 1., .1)),
    "CAS_BiSodCaseY_pente_borne_superbee_debar_2": (3, (5, 100, 5), (0, 0, 0), (.1, 1., .1)),
    "CAS_MonoPointtriple_pb_mixte_superbeeG_superbeeG": (2, (70, 30, 1), (0, 0, 0), (.07, .03, 0.)),
    "CAS_2D_Lag_SodCaseX_newton": (2, (100, 5, 1), (0, 0, 0), (1., .1, 0.)),
    "CAS_SodCaseX": (3, (100, 5, 5), (0, 0, 0), (1., .1, .1)),
    "CAS_MonoRiderTx": (3, (50, 50, 1), (0, 0, 0), (1., 1., .1)),
    "CAS_MonoRiderTy": (3, (50, 50, 1), (0, 0, 0), (1., 1., .1)),
    "CAS_MonoRiderT45": (3, (50, 50, 1), (0, 0, 0), (1., 1., .1)),
    "CAS_MonoRiderRotation": (3, (50, 50, 1), (0, 0, 0), (1., 1., .1)),
    "CAS_MonoRiderVortex": (3, (50, 50, 1), (0, 0, 0), (1., 1., .1)),
    "CAS_MonoRiderDeformation": (3, (50, 50, 1), (0, 0, 0), (1., 1., .1)),
    "CAS_MonoRiderVortexTimeReverse": (3, (50, 50, 1), (0, 0, 0), (1., 1., .1)),
    "CAS_MonoRiderDeformationTimeReverse": (3, (50, 50, 1), (0, 0, 0), (1., 1., .1)),
    "CAS_RiderTx": (3, (50, 50, 1), (0, 0, 0), (1., 1., .1)),
    "CAS_RiderTy": (3, (50, 50, 1), (0, 0, 0), (1., 1., .1)),
    "CAS_RiderT45": (3, (50, 50, 1), (0, 0, 0), (1., 1., .1)),
    "CAS_RiderRotation": (3, (50, 50, 1), (0, 0, 0), (1., 1., .1)),
    "CAS_RiderVortex": (3, (50, 50, 1), (0, 0, 0), (1., 1., .1)),
    "CAS_RiderDeformation": (3, (50, 50, 1), (0, 0, 0), (1., 1., .1)),
    "CAS_RiderVortexTimeReverse": (3, (50, 50, 1), (0, 0, 0), (1., 1., .1)),
    "CAS_RiderDeformationTimeReverse": (3, (50, 50, 1), (0, 0, 0), (1., 1., .1)),
    "CAS_MonoRiderTy_pb_simple_superbee": (3, (50, 50, 1), (0, 0, 0), (1., 1., .1)),
    "CAS_MonoRiderTy_pb_mixte_superbee_vanleer": (3, (50, 50, 1), (0, 0, 0), (1., 1., .1)),
    "CAS_MonoPointtriple_pb_sp_mixte_superbeeG_superbeeG": (2, (70, 30, 1), (0, 0, 0), (.07, .03, 0.)),
}
# Runs whose output holds no geometry: parts are the subdomains of the Cartesian generator
# in rank order (x fastest), each with its cells in local order (x fastest); their node
# fields are not kept (a part's node order is not recoverable without its geometry).
NSD_BLOCKS = {"CAS_Sedov": (2, 2, 2)}
# Lagrange-only runs on several subdomains: the final mesh has moved along the tube axis,
# so cells and nodes are placed by their rank along that axis within each line.
MOVING_AXIS = {"CAS_Lag_BiSodCaseZ": 2, "CAS_Lag_SodCaseX": 0, "CAS_2D_Lag_BiSodCaseX": 0,
               "CAS_2D_Lag_BiSodCaseY": 1, "CAS_2D_Lag_SodCaseY": 1}
TH_VARS = ["CellMass", "Pressure", "Density", "InternalEnergy"]
CELL_VARS = ["Density", "Pressure", "InternalEnergy", "CellMass", "PseudoViscosity", "Materiau"]


def _env_names(arc_path):
    arc = re.sub(r"<!--.*?-->", "", open(arc_path, encoding="latin-1").read(), flags=re.S)
    return re.findall(r"<environment>\s*<name>(\w+)</name>", arc)


def _read_geo(path):
    """Returns a list of parts: (node_xyz [n,3], connectivity [ne, k] 0-based)."""
    lines = open(path).read().split("\n")
    parts, i = [], 0
    while i < len(lines):
        if lines[i].strip() == "coordinates":
            n = int(lines[i + 1])
            vals = np.array([float(v) for v in lines[i + 2:i + 2 + 3 * n]])
            xyz = vals.reshape(3, n).T
            j = i + 2 + 3 * n
            etype = lines[j].strip()
            ne = int(lines[j + 1])
            conn = np.array([[int(t) for t in lines[j + 2 + e].split()] for e in range(ne)]) - 1
            parts.append((xyz, conn, etype))
            i = j + 2 + ne
        else:
            i += 1
    return parts


def _read_var(path, nper, ncomp=1):
    """Values per part; nper[p] = number of entries of part p."""
    lines = open(path).read().split("\n")
    out, i, p = [], 0, 0
    while i < len(lines):
        s = lines[i].strip()
        if s in ("quad4", "hexa8", "coordinates"):
            cnt = nper[p] * ncomp
            out.append(np.array([float(v) for v in lines[i + 1:i + 1 + cnt]]))
            i += 1 + cnt
            p += 1
        else:
            i += 1
    return out


def _block_maps(nsd, n):
    """Global cell / node ids of the parts of a 3D run on nsd subdomains without geometry."""
    cell_maps, node_maps = [], []
    b = [n[a] // nsd[a] for a in range(3)]
    for r in range(nsd[0] * nsd[1] * nsd[2]):
        p = (r % nsd[0], (r // nsd[0]) % nsd[1], r // (nsd[0] * nsd[1]))
        k, j, i = np.meshgrid(*[np.arange(p[a] * b[a], (p[a] + 1) * b[a]) for a in (2, 1, 0)], indexing="ij")
        cell_maps.append((i + n[0] * (j + n[1] * k)).ravel())
        k, j, i = np.meshgrid(*[np.arange(p[a] * b[a], (p[a] + 1) * b[a] + 1) for a in (2, 1, 0)], indexing="ij")
        node_maps.append((i + (n[0] + 1) * (j + (n[1] + 1) * k)).ravel())
    return cell_maps, node_maps


def _ranked_maps(parts, ax, dim, n, origin, dx):
    """Global cell / node ids of every part when the mesh moved along axis ax only: the
    other coordinates give the line, the rank of the coordinate along ax within the line
    (over all parts) gives the index along it."""
    nx, ny, nz = n[0], n[1], (n[2] if dim == 3 else 1)
    nn_ = [nx + 1, ny + 1, nz + 1]

    def ids(points, cells):
        pts = np.concatenate(points)
        owner = np.concatenate([np.full(len(p), q) for q, p in enumerate(points)])
        other = [a for a in range(dim) if a != ax]
        idx = np.zeros((len(pts), 3), dtype=int)
        for a in other:
            idx[:, a] = (np.floor((pts[:, a] - origin[a]) / dx[a]) if cells
                         else np.rint((pts[:, a] - origin[a]) / dx[a])).astype(int)
        key = idx[:, other[0]] + (10 ** 6) * (idx[:, other[1]] if len(other) > 1 else 0)
        order = np.lexsort((pts[:, ax], key))
        rank = np.empty(len(pts), dtype=int)
        sk = key[order]
        start = np.r_[0, np.flatnonzero(np.diff(sk)) + 1]
        line_start = np.repeat(start, np.diff(np.r_[start, len(sk)]))
        rank[order] = np.arange(len(pts)) - line_start
        idx[:, ax] = rank
        dims = [nx, ny, nz] if cells else nn_
        gid = idx[:, 0] + dims[0] * idx[:, 1] + (dims[0] * dims[1] * idx[:, 2] if dim == 3 else 0)
        return [gid[owner == q] for q in range(len(points))]

    cen = [xyz[conn].mean(axis=1) for xyz, conn, _ in parts]
    cell_maps = ids(cen, True)
    # nodes shared by two subdomains appear in both parts: rank the distinct positions
    allxyz = np.concatenate([xyz for xyz, _, _ in parts])
    uniq, inv = np.unique(np.round(allxyz, 12), axis=0, return_inverse=True)
    gu = ids([uniq], False)[0]
    sizes = np.cumsum([0] + [len(xyz) for xyz, _, _ in parts])
    node_maps = [gu[inv.ravel()[sizes[q]:sizes[q + 1]]] for q in range(len(parts))]
    return cell_maps, node_maps


def convert(root, case, outdir):
    dim, n, origin, length = CASES[case]
    nx, ny, nz = n
    if dim == 2:
        nz = 1
    d = os.path.join(root, case, "output")
    dep = os.path.join(d, "depouillement")
    res = {}
    names = _env_names(os.path.join(root, case, "Donnees.arc"))
    res["env_names"] = np.array(names)
    lines = open(os.path.join(dep, "ensight.case")).read().split()
    res["final_time"] = np.array(float(lines[-1]))
    # time histories
    times = None
    for k, en in enumerate(names):
        for v in TH_VARS:
            fn = os.path.join(d, "courbes", "gnuplot", f"{v}_{en}")
            if not os.path.exists(fn):
                continue
            a = np.atleast_2d(np.loadtxt(fn))
            res[f"th_{v}_{k}"] = a[:, 1]
            times = a[:, 0]
    res["th_time"] = times if times is not None else np.zeros(0)  # some runs wrote no history
    # geometry -> global index maps
    dx = [length[a] / n[a] if n[a] else 1. for a in range(3)]
    geo_dir = os.path.join(dep, "ensight.geo")
    if case in NSD_BLOCKS:
        cell_maps, node_maps = _block_maps(NSD_BLOCKS[case], (nx, ny, nz))
        parts = []
    else:
        parts = _read_geo(os.path.join(geo_dir, sorted(os.listdir(geo_dir))[-1]))
        cell_maps, node_maps = [], []
    single = len(parts) == 1 and not cell_maps  # sequential run: one part in global order (Lagrangian meshes move)
    if not single and case in MOVING_AXIS:
        cell_maps, node_maps = _ranked_maps(parts, MOVING_AXIS[case], dim, n, origin, dx)
        parts = []
    for xyz, conn, _ in parts:
        if single:
            cell_maps.append(np.arange(len(conn)))
            node_maps.append(np.arange(len(xyz)))
            continue
        cen = xyz[conn].mean(axis=1)
        ijk = [np.floor((cen[:, a] - origin[a]) / dx[a]).astype(int) for a in range(dim)]
        cid = ijk[0] + nx * ijk[1] + (nx * ny * ijk[2] if dim == 3 else 0)
        cell_maps.append(cid)
        nijk = [np.rint((xyz[:, a] - origin[a]) / dx[a]).astype(int) for a in range(dim)]
        nid = nijk[0] + (nx + 1) * nijk[1] + ((nx + 1) * (ny + 1) * nijk[2] if dim == 3 else 0)
        node_maps.append(nid)
    nc = nx * ny * nz
    nn = (nx + 1) * (ny + 1) * ((nz + 1) if dim == 3 else 1)
    for v in CELL_VARS:
        vd = os.path.join(dep, v)
        vals = _read_var(os.path.join(vd, sorted(os.listdir(vd))[-1]), [len(c) for c in cell_maps])
        g = np.full(nc, np.nan)
        for cm, pv in zip(cell_maps, vals):
            g[cm] = pv
        assert not np.isnan(g).any(), (case, v)
        res[f"cell_{v}"] = g
    if case in NSD_BLOCKS:  # node order within a part is not recoverable without geometry
        np.savez_compressed(os.path.join(outdir, case + ".npz"), **res)
        return
    vd = os.path.join(dep, "NodeMass")
    vals = _read_var(os.path.join(vd, sorted(os.listdir(vd))[-1]), [len(m) for m in node_maps])
    g = np.full(nn, np.nan)
    for nm, pv in zip(node_maps, vals):
        g[nm] = pv
    res["node_NodeMass"] = g
    vd = os.path.join(dep, "Velocity")
    if not os.path.isdir(vd):
        np.savez_compressed(os.path.join(outdir, case + ".npz"), **res)
        return
    vals = _read_var(os.path.join(vd, sorted(os.listdir(vd))[-1]), [len(m) for m in node_maps], 3)
    for c in range(3):
        g = np.full(nn, np.nan)
        for nm, pv in zip(node_maps, vals):
            g[nm] = pv.reshape(3, -1)[c]
        res[f"node_Velocity_{c}"] = g
    np.savez_compressed(os.path.join(outdir, case + ".npz"), **res)


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/NONREGRESSION"
    outdir = os.path.dirname(os.path.abspath(__file__))
    for c in sys.argv[2:] or CASES:
        convert(root, c, outdir)
        print("wrote", c)
