"""The CPU baseline's block matvec (oracle/hmat_cpu.py, the restatement of
blocks_process_MVM, src/Hmatrix.py:1445-1538) against the dense FP64 oracle
product, over pools built on the host exactly as the GPU build lays them out
(dense blocks from the oracle's traction operators, low-rank blocks from
truncated SVDs of the stacked 3m x 2n blocks)."""
import types

import numpy as np
import pytest
import torch

from oracle.hmat_cpu import CpuHMatrix
from pyquake3d_amd.cluster import build_block_partition, build_cluster_tree
from pyquake3d_amd.hmatrix import HMatrix, leaf_segments, leaf_tables, pack_u, plan_layout
from pyquake3d_amd.mesh import get_eleVec, undulating_fault

MU = LAM = 32e9


def host_hmatrix(oracle, eta, rel_cut):
    """An HMatrix-shaped object (pool, block tables, leaf tables) of a small
    undulating fault, and the six dense operators in cluster order."""
    mesh = undulating_fault(length=30e3, depth=10e3, nx=30, nz=10, seed=5)
    ev, xg = get_eleVec(mesh.nodelst, mesh.elelst)
    n = mesh.n
    t = build_cluster_tree(xg, leaf=32)
    P = build_block_partition(t, eta=eta)
    perm = t.perm
    K = oracle.traction_block(xg, ev, mesh.triangles(), perm, perm, MU, LAM)
    r0, m = t.start[P.rows], t.size[P.rows]
    c0, nn = t.start[P.cols], t.size[P.cols]
    lr_idx = np.where(P.adm)[0]
    ranks, fac = [], {}
    for bi in lr_idx:
        a, b = slice(r0[bi], r0[bi] + m[bi]), slice(c0[bi], c0[bi] + nn[bi])
        S = np.block([[K[0][a, b], K[3][a, b]], [K[1][a, b], K[4][a, b]],
                      [K[2][a, b], K[5][a, b]]])
        U, s, Vt = np.linalg.svd(S, full_matrices=False)
        k = int((s > rel_cut * s[0]).sum())
        ranks.append(k)
        fac[bi] = (U[:, :k] * s[:k], Vt[:k])
    L = plan_layout(r0, m, c0, nn, P.adm, lr_idx, np.array(ranks, dtype=np.int64))
    pool = np.zeros(max(L["total"], 1))
    for bi, base in zip(L["dense_idx"], L["d_base"]):
        bm, bn = m[bi], nn[bi]
        for op in range(6):
            blk = K[op][r0[bi]:r0[bi] + bm, c0[bi]:c0[bi] + bn]
            pool[base + op * bm * bn: base + (op + 1) * bm * bn] = blk.T.ravel()
    tiles, tor = leaf_tables(t)
    for bi, ub, vb in zip(L["lr_idx"], L["l_ub"], L["l_vb"]):
        U, V = fac[bi]
        pool[ub:ub + U.size] = pack_u(U, leaf_segments(tiles, tor, r0[bi], m[bi]))
        pool[vb:vb + V.size] = V.T.ravel()
    h = types.SimpleNamespace(
        n=n, pool=torch.from_numpy(pool), leaf_tiles=tiles, tile_of_row=tor,
        blocks=dict(r0=r0, m=m, c0=c0, n=nn, dense_idx=L["dense_idx"], d_base=L["d_base"],
                    lr_idx=L["lr_idx"], ranks=L["k"], l_ub=L["l_ub"], l_vb=L["l_vb"]))
    h.lowrank_factors = types.MethodType(HMatrix.lowrank_factors, h)
    return h, K


def _dense(K, v1, v2):
    return np.stack([K[0] @ v1 + K[3] @ v2, K[1] @ v1 + K[4] @ v2, K[2] @ v1 + K[5] @ v2])


def test_all_dense_blocks_equal_the_dense_product(oracle):
    """eta = 0: every block is near field, so the block loop is the dense
    FP64 product up to summation order."""
    h, K = host_hmatrix(oracle, eta=0.0, rel_cut=0.0)
    C = CpuHMatrix(h)
    assert not C.lowrank and len(C.dense) > 1
    rng = np.random.default_rng(1)
    v1, v2 = rng.standard_normal(h.n), rng.standard_normal(h.n)
    ref = _dense(K, v1, v2)
    y = C.matvec(v1, v2)
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()


def test_low_rank_blocks_within_their_truncation(oracle):
    """eta = 2 with SVD factors cut at 1e-10 of the largest singular value:
    the block loop (dense blocks plus U (V x) of the far field, scattered
    into the block rows) stays within 1e-8 of the dense product; a
    truncation at 1e-3 is visible (the low-rank path is really used), and
    the two-plane product gives the first two rows (to BLAS rounding: the
    row count changes the blocking of the products)."""
    rng = np.random.default_rng(2)
    v1, v2 = rng.standard_normal(900), rng.standard_normal(900)
    h, K = host_hmatrix(oracle, eta=2.0, rel_cut=1e-10)
    C = CpuHMatrix(h)
    assert C.lowrank and C.dense
    ref = _dense(K, v1[:h.n], v2[:h.n])
    y = C.matvec(v1[:h.n], v2[:h.n])
    assert np.abs(y - ref).max() <= 1e-8 * np.abs(ref).max()
    y2 = C.matvec(v1[:h.n], v2[:h.n], planes=2)
    assert np.abs(y2[:2] - y[:2]).max() <= 1e-13 * np.abs(y[:2]).max()
    assert not y2[2].any()
    hc, _ = host_hmatrix(oracle, eta=2.0, rel_cut=1e-3)
    yc = CpuHMatrix(hc).matvec(v1[:h.n], v2[:h.n])
    err = np.abs(yc - ref).max() / np.abs(ref).max()
    assert 1e-7 < err < 1e-1


@pytest.mark.parametrize("nproc", [1, 3])
def test_scattered_product_equals_single_process(oracle, nproc):
    """The block scatter over host processes (blocks dealt to ranks, partial
    products summed) gives the single-process product to rounding, for
    three and two row planes, on repeated calls."""
    from oracle.hmat_cpu import ScatterMatvec
    rng = np.random.default_rng(4)
    h, _ = host_hmatrix(oracle, eta=2.0, rel_cut=1e-6)
    C = CpuHMatrix(h)
    S = ScatterMatvec(C, nproc)
    try:
        for planes in (3, 2, 3):
            v1, v2 = rng.standard_normal(h.n), rng.standard_normal(h.n)
            ref = C.matvec(v1, v2, planes=planes)
            got = S.matvec(v1, v2, planes=planes)
            assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
            if planes == 2:
                assert not got[2].any()
    finally:
        S.close()
