"""The reference-shaped API (createHmatrix / BlockTree / ACA_worker /
blocks_process_MVM, src/Hmatrix.py) used the way the reference's driver uses
it (src/main.py:138-148, src/QDsim.py:906-917), against the dense oracle."""
import numpy as np
import pytest

from pyquake3d_amd.blocktree import createHmatrix
from pyquake3d_amd.mesh import get_eleVec, undulating_fault

pytestmark = pytest.mark.gpu
MU = LAM = 32e9


@pytest.fixture(scope="module")
def built(oracle):
    mesh = undulating_fault(length=30e3, depth=10e3, nx=36, nz=14, seed=6)
    eleVec, xg = get_eleVec(mesh.nodelst, mesh.elelst)
    tree = createHmatrix(xg, mesh.nodelst, mesh.elelst, eleVec, MU, LAM, False, mini_leaf=32)
    blocks = [tree.ACA_worker(b) for b in tree.blocks_to_process]
    n = mesh.n
    K = oracle.traction_block(xg, eleVec, mesh.triangles(), np.arange(n), np.arange(n), MU, LAM)
    return mesh, tree, blocks, dict(zip(("A1s", "A2s", "Bs", "A1d", "A2d", "Bd"), K))


def test_evaluators_match_oracle_columns_and_rows(gpu, pq3d_lib, built):
    mesh, tree, blocks, K = built
    rows = np.arange(0, mesh.n, 7)
    cols = np.arange(3, mesh.n, 5)
    A1s, A2s, Bs = tree.calc_stressfunc_fromCS(rows, cols)
    A1d, A2d, Bd = tree.calc_stressfunc_fromC_rowD(rows, cols)
    for name, got in zip(("A1s", "A2s", "Bs", "A1d", "A2d", "Bd"), (A1s, A2s, Bs, A1d, A2d, Bd)):
        ref = K[name][np.ix_(rows, cols)]
        assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max(), name
    a, b, c = tree.calc_rowsS(2, 5, rows, cols)
    assert np.array_equal(a, A1s[2:5])
    a, b, c = tree.calc_colsD(4, 9, rows, cols)
    assert np.array_equal(c, Bd[:, 4:9])


def test_get_ttstress_projection(gpu, pq3d_lib, built, oracle):
    mesh, tree, blocks, K = built
    tri = mesh.triangles()
    rows = np.arange(10)
    s, _ = oracle.stress_points(tree.xg[rows], tri[20, 0:3], tri[20, 3:6], tri[20, 6:9], 1.0, 0.0,
                                0.0, MU, LAM)
    Tra = tree.GetTtstress(s.T, rows)
    assert np.abs(Tra[:, 0] - K["A1s"][rows, 20]).max() <= 1e-10 * np.abs(K["A1s"][rows, 20]).max()


@pytest.mark.parametrize("op", ["A1s", "A2s", "Bs", "A1d", "A2d", "Bd"])
def test_blocks_process_mvm_each_operator(gpu, pq3d_lib, built, op):
    mesh, tree, blocks, K = built
    assert any(b.judaca for b in blocks) and any(not b.judaca for b in blocks)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(mesh.n)
    y = tree.blocks_process_MVM(x, blocks, op)
    ref = K[op] @ x
    scale = np.abs(K[op]).max() * np.abs(x).sum() / np.sqrt(mesh.n)
    assert np.abs(y - ref).max() <= 2e-3 * scale
