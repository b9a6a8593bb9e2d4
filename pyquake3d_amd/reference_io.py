"""Readers for the reference's model files, so that its example models run on
this build unchanged.

  read_msh_v2(path)        gmsh 2.2 ASCII mesh (src/readmsh.py:12-68)
  read_para(path)          "key: value" parameter file (src/QDsim.py:91-99)
  problem_from_parameters  a Problem (friction, stress, loading, initial
                           slip rate) from a parameter file and its mesh, as
                           QDsim.__init__ / Init_condition / read_parameter
                           set them up (src/QDsim.py:121-196, :298-306,
                           :387-655, :687-740)

Units follow the reference: coordinates in m, tractions in MPa (the 10-column
parameter file and the principal-stress values are taken as they are).
"""
from __future__ import annotations

import os

import numpy as np

from .mesh import FaultMesh, get_eleVec
from .problem import Problem


def read_msh_v2(path):
    """(nodelst (M, 3) float64 in m, elelst (N, 3) int64, 1-based).

    As the reference: node coordinates are multiplied by 1e3 (the files are
    in km), and every $Elements line with at least 8 fields contributes its
    last three integers (gmsh type-2 triangles with two tags); shorter lines
    (points, lines) are skipped."""
    nodes, eles = [], []
    section = None
    with open(path, "r") as f:
        for raw in f:
            line = raw.strip()
            if line.startswith("$"):
                section = {"$Nodes": "nodes", "$Elements": "elements"}.get(line)
                first = True
                continue
            if section is None:
                continue
            if first:            # the count line
                first = False
                continue
            parts = line.split()
            if section == "nodes" and len(parts) >= 4:
                nodes.append([float(parts[1]), float(parts[2]), float(parts[3])])
            elif section == "elements" and len(parts) >= 8:
                eles.append([int(parts[-3]), int(parts[-2]), int(parts[-1])])
    return (np.asarray(nodes, dtype=np.float64) * 1e3,
            np.asarray(eles, dtype=np.int64).reshape(-1, 3))


def read_para(path) -> dict:
    """Keys and values of every line holding ':' (text before the first ':'
    and between the first and second, both stripped), as the reference."""
    out = {}
    with open(path, "r") as f:
        for line in f:
            if ":" in line:
                parts = line.split(":")
                out[parts[0].strip()] = parts[1].strip()
    return out


def _flag(p, key):
    return p.get(key, "False") == "True"


def boundary_nodes(elelst):
    """1-based ids of the nodes on edges used by exactly one triangle
    (src/QDsim.py:56-80)."""
    e = np.asarray(elelst, dtype=np.int64)
    edges = np.concatenate([e[:, [0, 1]], e[:, [1, 2]], e[:, [2, 0]]])
    edges = np.sort(edges, axis=1)
    uniq, counts = np.unique(edges, axis=0, return_counts=True)
    return np.unique(uniq[counts == 1].ravel())


def _min_dist(points, targets, chunk=4096):
    out = np.empty(len(points))
    for s in range(0, len(points), chunk):
        d = points[s:s + chunk, None, :] - targets[None, :, :]
        out[s:s + chunk] = np.sqrt((d * d).sum(-1)).min(1)
    return out


def lab_edge_normal_stress(Tno, xg, nodelst, elelst, seed=42):
    """Normal stress of the laboratory model (Tn_edge, src/QDsim.py:309-377):
    within 2 mm of the mesh boundary away from the free surface it ramps
    linearly from 0.7 Tno[0] at the boundary to Tno[0]; elsewhere it is a
    smoothed random field: the ramped stress times (1 + U(-0.5, 0.5))
    (legacy numpy generator seeded with 42: a permutation of the cells, then
    the uniform draws), sampled by nearest neighbour on a 100 x 100 (x, y)
    grid, Gaussian-filtered (sigma 4 cells) and sampled back by nearest
    neighbour."""
    from scipy.interpolate import griddata
    from scipy.ndimage import gaussian_filter
    rs = np.random.RandomState(seed)
    Tno = np.array(Tno, dtype=np.float64)
    bn = boundary_nodes(elelst)
    bc = nodelst[bn - 1]
    bc_sb = bc[~(np.abs(bc[:, 2]) < 1e-5)]
    wedge = 0.1 * 2.0 * 1e-2
    ed, top = Tno[0] * 0.7, Tno[0]
    dis = _min_dist(xg, bc_sb) / wedge
    edge = dis < 1.0
    Tno[edge] = ed - (ed - top) * dis[edge]
    n = len(Tno)
    sample = rs.choice(np.arange(n), size=n, replace=False)
    arr = rs.uniform(low=-0.5, high=0.5, size=n)
    noisy = Tno.copy()
    noisy[sample] = Tno[sample] + Tno[sample] * arr
    xi = np.linspace(xg[:, 0].min(), xg[:, 0].max(), 100)
    yi = np.linspace(xg[:, 1].min(), xg[:, 1].max(), 100)
    Xi, Yi = np.meshgrid(xi, yi)
    Si = griddata((xg[:, 0], xg[:, 1]), noisy, (Xi, Yi), method="nearest")
    smooth = gaussian_filter(Si, sigma=4.0)
    val = griddata(np.stack((Xi.ravel(), Yi.ravel()), axis=1), smooth.ravel(), xg[:, [0, 1]],
                   method="nearest")
    val = np.where(np.isnan(val), Tno, val)
    return np.where(edge, Tno, val)


def problem_from_parameters(param_path, mesh_path=None):
    """Problem, half-space flag and parameter dict of a reference model.

    ``mesh_path`` defaults to the only ``*.msh`` file next to the parameter
    file. Supported: InputHetoparamter True (the 10-column file: rake, a, b,
    Dc, f0, -, Tno, V, shear and normal loading per element,
    src/QDsim.py:687-740) and False (principal stresses projected on each
    element, friction zones by distance to the mesh boundary, optional
    nucleation disc, src/QDsim.py:387-655), with the laboratory model's
    edge-tapered, smoothed random normal stress when Initlab is set
    (``lab_edge_normal_stress``)."""
    p = read_para(param_path)
    d = os.path.dirname(os.path.abspath(param_path))
    if mesh_path is None:
        msh = [f for f in sorted(os.listdir(d)) if f.endswith(".msh")]
        if len(msh) != 1:
            raise ValueError(f"expected one .msh file next to {param_path}, found {msh}")
        mesh_path = os.path.join(d, msh[0])
    nodelst, elelst = read_msh_v2(mesh_path)
    if not _flag(p, "Scale_km"):
        nodelst = nodelst / 1e3
    order = _flag(p, "Node_order")   # reversed receiver normals; pass on as jud_ele_order
    eleVec, xg = get_eleVec(nodelst, elelst, order)
    N = len(xg)
    mu = float(p["Shear modulus"])
    lam = float(p["Lame constants"])
    rho = float(p["Rock density"])
    V0 = float(p["Reference slip rate"])
    Vpl = float(p["Plate loading rate"])
    ev1, ev2, ev3 = eleVec[:, 0:3], eleVec[:, 3:6], eleVec[:, 6:9]

    # stresses projected on the element (src/QDsim.py:387-456)
    ssv_scale = float(p["Vertical principal stress"])
    ssh1_scale = float(p["Maximum horizontal principal stress"])
    ssv0_scale = float(p["Minimum horizontal principal stress"])
    trac_nor = float(p["Vertical principal stress value"])
    depth_var = _flag(p, "Vertical principal stress value varies with depth")
    if depth_var:
        turn = float(p["Turnning depth"])
        ssv = np.minimum(-xg[:, 2] / turn + 0.2, 1.0) * trac_nor * ssv_scale
    else:
        ssv = np.full(N, trac_nor * ssv_scale)
    ssh1, ssh2, ssv = -ssv * ssh1_scale, -ssv * ssv0_scale, -ssv
    ang = np.pi / 180.0 * float(p["Angle between ssh1 and X-axis"])
    c, s = np.cos(ang), np.sin(ang)
    s00 = c * c * ssh1 + s * s * ssh2
    s01 = c * s * ssh1 - c * s * ssh2
    s11 = s * s * ssh1 + c * c * ssh2
    tra = np.stack([s00 * ev3[:, 0] + s01 * ev3[:, 1], s01 * ev3[:, 0] + s11 * ev3[:, 1],
                    ssv * ev3[:, 2]], axis=1)
    Tt1o = (tra * ev1).sum(1)
    Tt2o = (tra * ev2).sum(1)
    solve_normal = _flag(p, "Normal traction solved from stress tensor")
    solve_shear = _flag(p, "Shear traction solved from stress tensor")
    Tno = (tra * ev3).sum(1) if solve_normal else ssv.copy()
    Tno = np.abs(Tno)
    Tt = np.sqrt(Tt1o * Tt1o + Tt2o * Tt2o)
    if _flag(p, "Rake solved from stress tensor"):
        rake = np.arctan2(Tt2o / Tt, Tt1o / Tt)
    else:
        rake = np.full(N, float(p["Fix_rake"])) / 180.0 * np.pi
    if not solve_normal and not depth_var:
        Tno = np.full(N, trac_nor)
    if _flag(p, "Initlab"):
        Tno = lab_edge_normal_stress(Tno, xg, nodelst, elelst)

    slipvC = np.full(N, Vpl)            # Grad_slpv_con(const=True)
    f0 = np.full(N, float(p["Reference friction coefficient"]))
    shear_loading = np.zeros(N)
    normal_loading = np.zeros(N)
    Tt_init = None
    if _flag(p, "InputHetoparamter"):
        vals = np.loadtxt(os.path.join(d, p["Inputparamter file"]), ndmin=2)[:N]
        rake, a, b, dc, f0 = vals[:, 0], vals[:, 1], vals[:, 2], vals[:, 3], vals[:, 4]
        Tno = vals[:, 6]
        V = vals[:, 7]
        shear_loading, normal_loading = vals[:, 8], vals[:, 9]
        Vpl_cell = slipvC
    else:
        bn = boundary_nodes(elelst)
        bc = nodelst[bn - 1]
        surf = np.abs(bc[:, 2]) < 1e-5
        bc_surface, bc_sb = bc[surf], bc[~surf]
        W = float(p["Widths of VS region"])
        Ws = float(p["Widths of surface VS region"])
        tr = float(p["Transition region from VS to VW region"])
        aVs = float(p["Rate-and-state parameters a in VS region"])
        bVs = float(p["Rate-and-state parameters b in VS region"])
        dcVs = float(p["Characteristic slip distance in VS region"])
        aVw = float(p["Rate-and-state parameters a in VW region"])
        bVw = float(p["Rate-and-state parameters b in VW region"])
        dcVw = float(p["Characteristic slip distance in VW region"])
        dist = _min_dist(xg, bc_sb)
        dis = dist / W
        dis_s = dis.copy()
        dis1 = (dist - W) / tr
        if len(bc_surface) > 10:                       # a free surface
            dist_s = _min_dist(xg, bc_surface)
            dis_s = dist_s / Ws
            dis1 = np.minimum(dist - W, dist_s - Ws) / tr
        vs = (dis < 1.0) | (dis_s < 1.0)
        trans = ~vs & (dis1 < 1.0)
        vw = ~vs & ~trans
        a = np.where(vs, aVs, np.where(vw, aVw, 0.0))
        b = np.where(vs, bVs, np.where(vw, bVw, 0.0))
        dc = np.where(vw, dcVw, dcVs)
        if _flag(p, "ChangefriA"):
            a = np.where(trans, aVs - (aVs - aVw) * dis1, a)
            b = np.where(trans, bVs, b)
        else:
            a = np.where(trans, aVs, a)
            b = np.where(trans, bVs - (bVs - bVw) * dis1, b)
        V = slipvC                      # the reference aliases these two arrays
        if _flag(p, "Set_nucleation"):
            loc = np.array([float(p["Nuclea_posx"]), float(p["Nuclea_posy"]),
                            float(p["Nuclea_posz"])])
            nuc = np.sqrt(((xg - loc) ** 2).sum(1)) < float(p["Radius of nucleation"])
            V = np.where(nuc, float(p["Initial slip rate in nucleation region"]), V)
            dc = np.where(nuc, float(p["Characteristic slip distance in nucleation region"]), dc)
            a = np.where(nuc, float(p["Rate-and-state parameters a in nucleation region"]), a)
            b = np.where(nuc, float(p["Rate-and-state parameters b in nucleation region"]), b)
        # slipv and slipvC are one array in the reference, so the nucleation
        # rate is also the plate rate of those cells (steady-state shear and
        # slip-rate deficit)
        Vpl_cell = V.copy()
        if solve_shear:
            Tt_init = Tt
    mesh = FaultMesh(nodelst, elelst, name=os.path.basename(mesh_path))
    prob = Problem(mesh, mu, lam, rho, V0, np.asarray(a, float), np.asarray(b, float),
                   np.asarray(dc, float), np.asarray(f0, float), np.asarray(Tno, float),
                   np.asarray(V, float).copy(), np.asarray(rake, float), np.asarray(Vpl_cell, float),
                   np.asarray(shear_loading, float), np.asarray(normal_loading, float),
                   fix_Tn=_flag(p, "Fix_Tn"), name=os.path.basename(d),
                   meta={"node_order": order, "Tt": Tt_init})
    return prob, _flag(p, "Half space"), p
