"""Readers of the reference's model files (gmsh 2.2 mesh, "key: value"
parameters, the 10-column element parameter file, the laboratory model's
normal stress) and the compact model archive. The example models are checked
against their committed archives when the reference tree is present."""
import os

import numpy as np
import pytest

from pyquake3d_amd.problem import load_problem, planar_c1, save_problem
from pyquake3d_amd.mesh import get_eleVec, planar_fault
from pyquake3d_amd.reference_io import (boundary_nodes, lab_edge_normal_stress,
                                        problem_from_parameters, read_msh_v2, read_para)

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
BP5 = "/root/reference/examples/BP5-QD/parameter.txt"
EXAMPLES = "/root/reference/examples"
needs_reference = pytest.mark.skipif(not os.path.isdir(EXAMPLES),
                                     reason="the reference's example files are not here")

MSH = """$MeshFormat\r
2.2 0 8\r
$EndMeshFormat\r
$Nodes\r
4\r
1 0 0 0\r
2 1 0 0\r
3 0 0 -1\r
4 1 0 -1.5\r
$EndNodes\r
$Elements\r
4\r
1 15 2 0 1 1\r
2 1 2 0 1 1 2\r
3 2 2 0 1 1 3 2\r
4 2 2 0 1 2 3 4\r
$EndElements\r
"""


def test_msh_reader_scales_km_and_keeps_triangles(tmp_path):
    f = tmp_path / "m.msh"
    f.write_bytes(MSH.encode())
    nodes, eles = read_msh_v2(str(f))
    assert np.array_equal(nodes, 1e3 * np.array([[0, 0, 0], [1, 0, 0], [0, 0, -1], [1, 0, -1.5]]))
    # the point (6 fields) and the line (7 fields) are skipped
    assert np.array_equal(eles, [[1, 3, 2], [2, 3, 4]])


def test_parameter_reader_takes_text_between_first_colons(tmp_path):
    f = tmp_path / "p.txt"
    f.write_text("#Files\nShear modulus: 3e10\nA: b: c\nno colon here\nFix_Tn:False\n")
    p = read_para(str(f))
    assert p == {"Shear modulus": "3e10", "A": "b", "Fix_Tn": "False"}


def test_boundary_nodes_of_two_triangles():
    assert np.array_equal(boundary_nodes(np.array([[1, 2, 3], [2, 4, 3]])), [1, 2, 3, 4])
    # a square split in four around a centre node: the centre is interior
    t = np.array([[1, 2, 5], [2, 3, 5], [3, 4, 5], [4, 1, 5]])
    assert np.array_equal(boundary_nodes(t), [1, 2, 3, 4])


def test_model_archive_round_trip(tmp_path):
    p = planar_c1(nx=6, nz=3)
    f = tmp_path / "m.npz"
    save_problem(str(f), p, halfspace=True)
    q, hs = load_problem(str(f))
    assert hs and q.n == p.n and q.fix_Tn == p.fix_Tn
    for k in ("a", "b", "dc", "f0", "Tno", "slipv", "rake", "Vpl"):
        assert np.array_equal(getattr(q, k), getattr(p, k)), k
    for x, y in zip(q.initial_state(), p.initial_state()):
        assert np.array_equal(x, y)


@pytest.mark.skipif(not os.path.exists(BP5), reason="reference example tree not present")
@needs_reference
def test_bp5_example_matches_its_archive():
    """The reader on the reference's BP5-QD files (9,250 triangles, half
    space, 10-column element file) equals the committed archive; the initial
    state is the steady state at Vpl except the 3e-2 m/s nucleation disc."""
    p, hs, para = problem_from_parameters(BP5)
    q, hs2 = load_problem(os.path.join(GOLDEN, "bp5_example.npz"))
    assert hs and hs2 and p.n == q.n == 9250
    assert np.array_equal(p.mesh.elelst, q.mesh.elelst)
    assert np.array_equal(p.mesh.nodelst, q.mesh.nodelst)
    for k in ("a", "b", "dc", "f0", "Tno", "slipv", "rake", "Vpl", "shear_loading",
              "normal_loading"):
        assert np.array_equal(getattr(p, k), getattr(q, k)), k
    assert set(np.unique(q.slipv)) == {1e-9, 3e-2}
    Tno, Tt1o, Tt2o, state, *_ = q.initial_state()
    slow = q.slipv == 1e-9
    # steady state at V = Vpl: state = f0 + b ln(V0/Vpl)
    assert np.allclose(state[slow], (q.f0 + q.b * np.log(q.V0 / q.Vpl))[slow], rtol=1e-12)


@needs_reference
@pytest.mark.parametrize("d,name,n,hs", [("Lab-model", "lab_example.npz", 14336, False),
                                         ("EAFZ-model", "eafz_example.npz", 41800, True)])
def test_examples_match_their_archives(d, name, n, hs):
    p, h, para = problem_from_parameters(os.path.join(EXAMPLES, d, "parameter.txt"))
    q, h2 = load_problem(os.path.join(GOLDEN, name))
    assert h == h2 == hs and p.n == q.n == n
    assert np.array_equal(p.mesh.elelst, q.mesh.elelst)
    assert np.array_equal(p.mesh.nodelst, q.mesh.nodelst)
    for k in ("a", "b", "dc", "f0", "Tno", "slipv", "rake", "Vpl"):
        assert np.array_equal(getattr(p, k), getattr(q, k)), k


def _dipping_lab_mesh():
    """A 4 x 2 cm fault dipping 30 degrees, its top edge at z = 0."""
    m = planar_fault(length=0.04, depth=0.02, nx=24, nz=12)
    nodes = m.nodelst.copy()
    nodes[:, 1] = -nodes[:, 2] / np.tan(np.radians(30.0))
    return nodes, m.elelst


def test_lab_normal_stress_edge_taper_and_smoothing():
    """Within 2 mm of the boundary (the free-surface edge excluded) the
    stress ramps from 0.7 Tno[0] to Tno[0] with the distance; elsewhere it
    is the random field smoothed: inside the range of the unsmoothed field
    and far less spread than it; fixed by the seed."""
    nodes, ele = _dipping_lab_mesh()
    ev, xg = get_eleVec(nodes, ele)
    T0 = np.full(len(xg), 60.0)
    out = lab_edge_normal_stress(T0, xg, nodes, ele)
    bn = boundary_nodes(ele)
    bc = nodes[bn - 1]
    bc = bc[np.abs(bc[:, 2]) >= 1e-5]
    dist = np.sqrt(((xg[:, None, :] - bc[None, :, :]) ** 2).sum(-1)).min(1)
    edge = dist < 0.002
    assert 0 < edge.sum() < len(xg)
    ramp = 42.0 + 18.0 * dist / 0.002
    assert np.allclose(out[edge], ramp[edge], rtol=0, atol=1e-12)
    inner = out[~edge]
    assert inner.min() >= 0.5 * 42.0 and inner.max() <= 1.5 * 60.0
    # the unsmoothed field has the spread of U(-0.5, 0.5) (std 0.29 of the mean)
    assert inner.std() < 0.1 * inner.mean()
    assert np.array_equal(out, lab_edge_normal_stress(T0, xg, nodes, ele))
    assert not np.array_equal(out, lab_edge_normal_stress(T0, xg, nodes, ele, seed=7))


def test_legacy_seeded_draws_equal_a_seeded_generator_object():
    """The reference seeds numpy's global generator with 42, then draws a
    permutation (choice without replacement) and uniforms; a RandomState(42)
    object gives the same sequence."""
    n = 1000
    np.random.seed(42)
    s1 = np.random.choice(np.arange(n), size=n, replace=False)
    a1 = np.random.uniform(low=-0.5, high=0.5, size=n)
    rs = np.random.RandomState(42)
    assert np.array_equal(s1, rs.choice(np.arange(n), size=n, replace=False))
    assert np.array_equal(a1, rs.uniform(low=-0.5, high=0.5, size=n))
