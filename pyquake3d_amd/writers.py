"""Output files of the reference, so that its post-processing reads this
build's results unchanged.

  write_vtk(path, nodelst, elelst, cells)  legacy-ASCII VTK of the fault with
                                           the reference's cell fields, names
                                           and number formats (QDsim.ouputVTK,
                                           src/QDsim.py:1142-1289)
  vtk_cells(snapshot, problem)             those cell fields from a simulator
                                           snapshot (original element order)
  read_vtk(path)                           nodes, elements and cell fields of
                                           such a file (the fields the
                                           reference's restart reader takes
                                           back, src/QDsim.py:656-682), without
                                           pyvista
  restart_fields(cells)                    the simulator state those fields
                                           hold (rake in radians)
  element_file_from_vtk(vtk, out)          the 10-column element file made from
                                           a run's VTK output, as
                                           examples/HF-model/create_init/
                                           initcond.py makes it, without pyvista
  model_summary(problem)                   element sizes, nucleation size and
                                           cohesive zone (src/QDsim.py:258-295)
  StateLog                                 state.txt: header, one line per
                                           step, run time (src/main.py:83-99,
                                           :177, :232-236)

Units as the reference: tractions in MPa, slip rates in m/s, slip in m.
"""
from __future__ import annotations

import datetime
import math

import numpy as np

# (field name, printf format) in the reference's order (src/QDsim.py:1165-1274)
VTK_FIELDS = (
    ("Normal_[MPa]", "%f"),
    ("Shear_[MPa]", "%f"),
    ("Shear_1[MPa]", "%f"),
    ("Shear_2[MPa]", "%f"),
    ("rake[Degree]", "%f"),
    ("state", "%f"),
    ("Slipv[m/s]", "%.40f"),
    ("Slipv1[m/s]", "%.40f"),
    ("Slipv2[m/s]", "%.40f"),
    ("a", "%f"),
    ("b", "%f"),
    ("a-b", "%f"),
    ("dc", "%.10f"),
    ("fric", "%f"),
    ("slip", "%.20f"),
    ("slip1", "%.20f"),
    ("slip2", "%.20f"),
    ("slip_plate", "%.15f"),
)


def _row(fmt, values) -> str:
    return "".join((fmt % v) + " " for v in np.asarray(values, dtype=np.float64).tolist())


def write_vtk(path, nodelst, elelst, cells: dict) -> None:
    """Legacy-ASCII unstructured grid of triangles (nodes in m, ``elelst``
    1-based) with one SCALARS block per entry of VTK_FIELDS, each taken from
    ``cells`` (length N, original element order), written with the
    reference's formats, spacing and line breaks."""
    nodes = np.asarray(nodelst, dtype=np.float64)
    ele = np.asarray(elelst, dtype=np.int64)
    ne = ele.shape[0]
    missing = [name for name, _ in VTK_FIELDS if name not in cells]
    if missing:
        raise KeyError(f"cell fields missing: {missing}")
    for name, _ in VTK_FIELDS:
        if np.shape(cells[name]) != (ne,):
            raise ValueError(f"cell field {name} must have one value per element")
    parts = ["# vtk DataFile Version 3.0\n", "test\n", "ASCII\n", "DATASET  UNSTRUCTURED_GRID\n",
             f"POINTS {nodes.shape[0]} float\n"]
    parts.extend("%f %f %f\n" % (x, y, z) for x, y, z in nodes.tolist())
    parts.append(f"CELLS {ne} {ne * 4}\n")
    parts.extend("3 %d %d %d\n" % (a - 1, b - 1, c - 1) for a, b, c in ele.tolist())
    parts.append(f"CELL_TYPES {ne}\n")
    parts.append("5 " * ne + "\n")
    parts.append("CELL_DATA %d " % ne)
    for name, fmt in VTK_FIELDS:
        parts.append(f"SCALARS {name} float\nLOOKUP_TABLE default\n")
        parts.append(_row(fmt, cells[name]) + "\n")
    with open(path, "w") as f:
        f.write("".join(parts))


def vtk_cells(snap: dict, problem) -> dict:
    """The reference's VTK cell fields from ``QDsim.snapshot()`` (or the same
    keys from ``EnsembleQDsim.instance``) and the problem's friction
    parameters: Tt = |(Tt1o, Tt2o)|, fric = Tt / Tno, slip = |(slip1, slip2)|
    (src/QDsim.py:971, :990, :993)."""
    Tt = np.sqrt(snap["Tt1o"] ** 2 + snap["Tt2o"] ** 2)
    a, b = np.asarray(problem.a, float), np.asarray(problem.b, float)
    return {
        "Normal_[MPa]": snap["Tno"], "Shear_[MPa]": Tt, "Shear_1[MPa]": snap["Tt1o"],
        "Shear_2[MPa]": snap["Tt2o"], "rake[Degree]": snap["rake"] * 180.0 / np.pi,
        "state": snap["state"], "Slipv[m/s]": snap["slipv"], "Slipv1[m/s]": snap["slipv1"],
        "Slipv2[m/s]": snap["slipv2"], "a": a, "b": b, "a-b": a - b,
        "dc": np.asarray(problem.dc, float), "fric": Tt / snap["Tno"],
        "slip": np.sqrt(snap["slip1"] ** 2 + snap["slip2"] ** 2), "slip1": snap["slip1"],
        "slip2": snap["slip2"], "slip_plate": np.asarray(problem.Vpl, float),
    }


def read_vtk(path) -> dict:
    """{"nodes": (M, 3), "elelst": (N, 3) 1-based, "cells": {name: (N,)}} of
    a legacy-ASCII unstructured grid of triangles with SCALARS cell data."""
    with open(path, "r") as f:
        tok = f.read().split()
    i = 0

    def find(word):
        nonlocal i
        while tok[i] != word:
            i += 1
        i += 1

    find("POINTS")
    m = int(tok[i])
    i += 2
    nodes = np.asarray(tok[i:i + 3 * m], dtype=np.float64).reshape(m, 3)
    i += 3 * m
    find("CELLS")
    n = int(tok[i])
    i += 2
    cl = np.asarray(tok[i:i + 4 * n], dtype=np.int64).reshape(n, 4)
    if (cl[:, 0] != 3).any():
        raise ValueError(f"{path}: only triangle cells are supported")
    i += 4 * n
    find("CELL_DATA")
    if int(tok[i]) != n:
        raise ValueError(f"{path}: CELL_DATA count differs from CELLS")
    i += 1
    cells = {}
    while i < len(tok):
        if tok[i] != "SCALARS":
            raise ValueError(f"{path}: unexpected token {tok[i]!r} in cell data")
        name = tok[i + 1]
        i += 3                                   # SCALARS name type
        if tok[i] == "LOOKUP_TABLE":
            i += 2
        cells[name] = np.asarray(tok[i:i + n], dtype=np.float64)
        i += n
    return {"nodes": nodes, "elelst": cl[:, 1:] + 1, "cells": cells}


def restart_fields(cells: dict) -> dict:
    """Simulator state held by a VTK file, keyed as ``QDsim.snapshot()``:
    the fields the reference's read_vtk takes back (src/QDsim.py:660-682),
    rake converted to radians."""
    return {
        "Tno": cells["Normal_[MPa]"], "Tt1o": cells["Shear_1[MPa]"],
        "Tt2o": cells["Shear_2[MPa]"], "state": cells["state"], "slipv": cells["Slipv[m/s]"],
        "slipv1": cells["Slipv1[m/s]"], "slipv2": cells["Slipv2[m/s]"],
        "rake": cells["rake[Degree]"] / 180.0 * np.pi, "slip1": cells["slip1"],
        "slip2": cells["slip2"],
    }


def element_file_from_vtk(vtk_path, out_path, f0=0.6) -> int:
    """Write the 10-column element file (rake, a, b, Dc, f0, shear traction,
    normal traction, slip rate, shear loading 0, normal loading 0) that
    examples/HF-model/create_init/initcond.py writes from a run's VTK output,
    with its formats; returns the number of rows. As in that script, the
    rake column holds the VTK's degrees, while the element-file reader takes
    it in radians (src/QDsim.py:698, :729-730)."""
    c = read_vtk(vtk_path)["cells"]
    rows = zip(c["rake[Degree]"].tolist(), c["a"].tolist(), c["b"].tolist(), c["dc"].tolist(),
               c["Shear_[MPa]"].tolist(), c["Normal_[MPa]"].tolist(),
               c["Slipv[m/s]"].tolist())
    n = 0
    with open(out_path, "w") as f:
        for rake, a, b, dc, tt, tn, v in rows:
            f.write("%f %f %f %f %g %f %f %.20f 0 0\n" % (rake, a, b, dc, f0, tt, tn, v))
            n += 1
    return n


def model_summary(problem) -> dict:
    """Largest and mean element size (longest side), critical nucleation size
    hRA and cohesive zone A0 from max b, min a, min Dc and the mean normal
    stress (src/QDsim.py:258-295)."""
    nodes, ele = problem.mesh.nodelst, problem.mesh.elelst - 1
    P1, P2, P3 = nodes[ele[:, 0]], nodes[ele[:, 1]], nodes[ele[:, 2]]
    side = np.stack([np.linalg.norm(P1 - P2, axis=1), np.linalg.norm(P1 - P3, axis=1),
                     np.linalg.norm(P2 - P3, axis=1)], 1).max(1)
    b = float(np.max(problem.b))
    a = float(np.min(problem.a))
    L = float(np.min(problem.dc))
    sigma = float(np.mean(np.asarray(problem.Tno) * 1e6))
    mu = float(problem.mu)
    hRA = 2.0 / math.pi * mu * b * L / (b - a) / (b - a) / sigma * math.pi * math.pi / 4.0
    A0 = 9.0 * math.pi / 32 * mu * L / (b * sigma)
    return {"maxsize": float(side.max()), "ave_elesize": float(side.mean()), "hRA": hRA,
            "A0": A0}


class StateLog:
    """state.txt of a run: the header (model sizes, elastic constants,
    element and nucleation sizes), one line per accepted step (step, h,
    max |V1|, max |V2|, time in s, time in days) and the run time, with the
    reference's formats (src/main.py:83-99, :177, :232-236)."""

    def __init__(self, path, problem, *, mesh_file="", param_file="", start=None):
        p = problem
        mu, lam = float(p.mu), float(p.lambda_)
        s = model_summary(p)
        self.f = open(path, "w")
        w = self.f.write
        w("Program start time: %s\n" % str(start or datetime.datetime.now()))
        w("Input msh geometry file:%s\n" % mesh_file)
        w("Input parameter file:%s\n" % param_file)
        w("Number of Node:%d\n" % p.mesh.nodelst.shape[0])
        w("Number of Element:%d\n" % p.mesh.elelst.shape[0])
        w("Cs:%f\n" % float(p.Cs))
        w("First Lamé constants:%f\n" % lam)
        w("Shear Modulus:%f\n" % mu)
        w("Youngs Modulus:%f\n" % (mu * (3.0 * lam + 2.0 * mu) / (lam + mu)))
        w("Poissons ratio:%f\n" % (lam / 2.0 / (lam + mu)))
        w("maximum element size:%f\n" % s["maxsize"])
        w("average elesize:%f\n" % s["ave_elesize"])
        w("Critical nucleation size:%f\n" % s["hRA"])
        w("Cohesive zone::%f\n" % s["A0"])
        w("iteration time_step(s) maximum_slip_rate(m/s) time(s) time(h)\n")

    def step(self, i, h, max_v1, max_v2, time) -> None:
        self.f.write("%d %f %f %.16e %f %f\n" % (i, h, max_v1, max_v2, time,
                                                 time / 3600.0 / 24.0))

    def close(self, seconds, end=None) -> None:
        self.f.write("Program end time: %s\n" % str(end or datetime.datetime.now()))
        self.f.write("Time taken: %.2f seconds\n" % seconds)
        self.f.close()
