"""The reference's output files (VTK cell data, state.txt): layout, field names
and number formats as src/QDsim.py:1142-1289 and src/main.py:83-99, :177,
:232-236 write them; VTK round trips through the reader; the element-size and
nucleation-size summary against closed forms."""
import math

import numpy as np
import pytest

from pyquake3d_amd import writers
from pyquake3d_amd.mesh import FaultMesh
from pyquake3d_amd.problem import planar_c1

# the fields the reference's restart reader takes back (src/QDsim.py:660-682)
RESTART_NAMES = {"rake[Degree]", "Shear_1[MPa]", "Shear_2[MPa]", "Shear_[MPa]", "Normal_[MPa]",
                 "Slipv1[m/s]", "Slipv2[m/s]", "Slipv[m/s]", "slip", "slip1", "slip2", "state",
                 "fric"}


def _snapshot(n, rng):
    return dict(Tno=rng.uniform(20, 30, n), Tt1o=rng.uniform(-5, 15, n),
                Tt2o=rng.uniform(-3, 3, n), state=rng.uniform(0.5, 0.8, n),
                slipv1=10 ** rng.uniform(-12, 0, n), slipv2=10 ** rng.uniform(-14, -2, n),
                slipv=10 ** rng.uniform(-12, 0, n), rake=rng.uniform(-np.pi, np.pi, n),
                slip1=rng.uniform(0, 10, n), slip2=rng.uniform(-1, 1, n))


@pytest.fixture(scope="module")
def small():
    return planar_c1(nx=6, nz=4)


def test_vtk_layout_matches_reference_writer(tmp_path, small):
    p = small
    snap = _snapshot(p.n, np.random.default_rng(0))
    path = tmp_path / "step0.vtk"
    writers.write_vtk(path, p.mesh.nodelst, p.mesh.elelst, writers.vtk_cells(snap, p))
    lines = path.read_text().split("\n")
    assert lines[:5] == ["# vtk DataFile Version 3.0", "test", "ASCII",
                         "DATASET  UNSTRUCTURED_GRID", f"POINTS {len(p.mesh.nodelst)} float"]
    m, n = len(p.mesh.nodelst), p.n
    x, y, z = p.mesh.nodelst[1]
    assert lines[6] == "%f %f %f" % (x, y, z)
    assert lines[5 + m] == f"CELLS {n} {4 * n}"
    a, b, c = p.mesh.elelst[0] - 1
    assert lines[6 + m] == f"3 {a} {b} {c}"                 # 0-based node ids
    assert lines[6 + m + n] == f"CELL_TYPES {n}"
    assert lines[7 + m + n] == "5 " * n
    # "CELL_DATA n " and the first SCALARS header share one line
    assert lines[8 + m + n] == f"CELL_DATA {n} SCALARS Normal_[MPa] float"
    assert lines[9 + m + n] == "LOOKUP_TABLE default"
    assert lines[10 + m + n] == "".join("%f " % v for v in snap["Tno"])
    names = [ln.split()[1] for ln in lines if "SCALARS" in ln and not ln.startswith("CELL_DATA")]
    names = ["Normal_[MPa]"] + names
    assert names == [nm for nm, _ in writers.VTK_FIELDS]
    assert RESTART_NAMES <= set(names)
    # number formats: 40 decimals for slip rates, 20 for slip, 10 for Dc
    i = lines.index("SCALARS Slipv[m/s] float")
    assert lines[i + 2].split()[0] == "%.40f" % snap["slipv"][0]
    i = lines.index("SCALARS slip float")
    assert lines[i + 2].split()[0] == "%.20f" % math.hypot(snap["slip1"][0], snap["slip2"][0])
    i = lines.index("SCALARS dc float")
    assert lines[i + 2].split()[0] == "%.10f" % p.dc[0]
    assert lines[-1] == ""


def test_vtk_round_trip_and_restart_fields(tmp_path, small):
    p = small
    snap = _snapshot(p.n, np.random.default_rng(1))
    cells = writers.vtk_cells(snap, p)
    path = tmp_path / "s.vtk"
    writers.write_vtk(path, p.mesh.nodelst, p.mesh.elelst, cells)
    got = writers.read_vtk(path)
    assert np.array_equal(got["elelst"], p.mesh.elelst)
    assert np.abs(got["nodes"] - p.mesh.nodelst).max() <= 5e-7
    for name, fmt in writers.VTK_FIELDS:
        digits = int(fmt[2:-1]) if fmt != "%f" else 6
        assert np.abs(got["cells"][name] - cells[name]).max() <= 0.51 * 10.0 ** -digits, name
    # fric = Tt / Tno and Tt = |(Tt1o, Tt2o)| (src/QDsim.py:971, :993)
    Tt = np.hypot(snap["Tt1o"], snap["Tt2o"])
    assert np.allclose(cells["fric"], Tt / snap["Tno"])
    assert np.allclose(cells["Shear_[MPa]"], Tt)
    assert np.allclose(cells["a-b"], p.a - p.b)
    r = writers.restart_fields(got["cells"])
    assert set(r) == set(snap)
    assert np.abs(r["rake"] - snap["rake"]).max() <= 1e-6 * np.pi / 180 * 1.01
    assert np.abs(r["slipv"] - snap["slipv"]).max() <= 1e-40 + 1e-16 * snap["slipv"].max()


def test_vtk_rejects_missing_or_short_fields(tmp_path, small):
    p = small
    cells = writers.vtk_cells(_snapshot(p.n, np.random.default_rng(2)), p)
    bad = dict(cells)
    del bad["fric"]
    with pytest.raises(KeyError):
        writers.write_vtk(tmp_path / "x.vtk", p.mesh.nodelst, p.mesh.elelst, bad)
    bad = dict(cells)
    bad["a"] = bad["a"][:-1]
    with pytest.raises(ValueError):
        writers.write_vtk(tmp_path / "x.vtk", p.mesh.nodelst, p.mesh.elelst, bad)


def test_model_summary_closed_form(small):
    # one right triangle with legs 3 and 4 m: longest side 5
    mesh = FaultMesh(np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0], [0.0, 0.0, -4.0]]),
                     np.array([[1, 2, 3]]))
    p = small
    q = type(p)(mesh, p.mu, p.lambda_, p.density, p.V0, np.array([0.01]), np.array([0.015]),
                np.array([0.2]), np.array([0.6]), np.array([50.0]), np.array([1e-9]),
                np.array([0.0]), np.array([1e-9]))
    s = writers.model_summary(q)
    assert s["maxsize"] == 5.0 and s["ave_elesize"] == 5.0
    sigma = 50e6
    hRA = 2 / np.pi * p.mu * 0.015 * 0.2 / 0.005 ** 2 / sigma * np.pi ** 2 / 4
    assert s["hRA"] == pytest.approx(hRA, rel=1e-14)
    assert s["A0"] == pytest.approx(9 * np.pi / 32 * p.mu * 0.2 / (0.015 * sigma), rel=1e-14)


def test_state_log_lines(tmp_path, small):
    path = tmp_path / "state.txt"
    log = writers.StateLog(path, small, mesh_file="m.msh", param_file="parameter.txt",
                           start="T0")
    log.step(3, 0.5, 1e-9, 2e-9, 86400.0)
    log.close(12.345, end="T1")
    lines = path.read_text().split("\n")
    assert lines[0] == "Program start time: T0"
    assert lines[1] == "Input msh geometry file:m.msh"
    assert lines[3] == f"Number of Node:{len(small.mesh.nodelst)}"
    assert lines[4] == f"Number of Element:{small.n}"
    assert lines[5] == "Cs:%f" % math.sqrt(small.mu / small.density)
    assert lines[9] == "Poissons ratio:%f" % (small.lambda_ / 2 / (small.lambda_ + small.mu))
    assert lines[14] == "iteration time_step(s) maximum_slip_rate(m/s) time(s) time(h)"
    assert lines[15] == "3 0.500000 0.000000 2.0000000000000001e-09 86400.000000 1.000000"
    assert lines[16] == "Program end time: T1"
    assert lines[17] == "Time taken: 12.35 seconds"


def test_element_file_from_vtk(tmp_path, small):
    """The HF-model workflow: a run's VTK output -> 10-column element file
    with the columns and formats of examples/HF-model/create_init/initcond.py."""
    p = small
    snap = _snapshot(p.n, np.random.default_rng(3))
    cells = writers.vtk_cells(snap, p)
    writers.write_vtk(tmp_path / "step0.vtk", p.mesh.nodelst, p.mesh.elelst, cells)
    n = writers.element_file_from_vtk(tmp_path / "step0.vtk", tmp_path / "initcondion.txt")
    assert n == p.n
    lines = (tmp_path / "initcondion.txt").read_text().splitlines()
    f = lines[0].split()
    assert len(f) == 10 and f[4] == "0.6" and f[8] == "0" and f[9] == "0"
    assert f[0] == "%f" % float("%f" % (snap["rake"][0] * 180 / np.pi))
    assert f[7] == "%.20f" % float("%.40f" % snap["slipv"][0])
    vals = np.loadtxt(tmp_path / "initcondion.txt")
    assert np.allclose(vals[:, 1], p.a, atol=5e-7) and np.allclose(vals[:, 6], snap["Tno"], atol=1e-6)
    assert np.allclose(vals[:, 5], np.hypot(snap["Tt1o"], snap["Tt2o"]), atol=1e-6)
