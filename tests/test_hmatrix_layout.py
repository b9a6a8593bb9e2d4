"""Host logic of the H-matrix on the CPU: cluster tree, block partition,
pool layout and matvec work lists. The two matvec kernels are emulated in
numpy item by item (same indexing as pyquake3d_amd/csrc/hmat_kernels.hip) over
a pool filled from the dense FP64 oracle (near field) and truncated SVDs of the
stacked 3m x 2n blocks (far field), and compared with the dense product."""
import numpy as np
import pytest
import torch

from pyquake3d_amd.cluster import build_block_partition, build_cluster_tree
from pyquake3d_amd.hmatrix import (build_worklists, leaf_segments, leaf_tables, pack_u,
                                   plan_layout, unpack_u)
from pyquake3d_amd.mesh import get_eleVec, undulating_fault

MU = LAM = 32e9


def test_cluster_tree_partitions_and_is_contiguous():
    mesh = undulating_fault(length=30e3, depth=10e3, nx=40, nz=15, seed=3)
    _, xg = get_eleVec(mesh.nodelst, mesh.elelst)
    t = build_cluster_tree(xg, leaf=32)
    assert sorted(t.perm.tolist()) == list(range(mesh.n))
    leaves = t.leaves()
    assert (t.size[leaves] <= 32).all()
    # leaves tile [0, n) without gaps
    s = t.start[leaves]
    assert s[0] == 0 and (np.diff(s) == t.size[leaves][:-1]).all()
    assert s[-1] + t.size[leaves][-1] == mesh.n
    # every internal node is split at the bbox midpoint of its longest extent
    for k in range(t.n_nodes):
        if t.left[k] < 0:
            continue
        pts = xg[t.perm[t.start[k]:t.start[k] + t.size[k]]]
        d = int(np.argmax(pts.max(0) - pts.min(0)))
        g = (pts[:, d].min() + pts[:, d].max()) / 2
        L = xg[t.perm[t.start[t.left[k]]:t.start[t.left[k]] + t.size[t.left[k]]]]
        R = xg[t.perm[t.start[t.right[k]]:t.start[t.right[k]] + t.size[t.right[k]]]]
        assert (L[:, d] <= g).all() and (R[:, d] > g).all()


def test_block_partition_covers_matrix_once():
    mesh = undulating_fault(length=30e3, depth=10e3, nx=40, nz=15, seed=3)
    _, xg = get_eleVec(mesh.nodelst, mesh.elelst)
    t = build_cluster_tree(xg, leaf=32)
    P = build_block_partition(t, eta=2.0)
    cover = np.zeros((mesh.n, mesh.n), dtype=np.int32)
    for a, b in zip(P.rows, P.cols):
        cover[t.start[a]:t.start[a] + t.size[a], t.start[b]:t.start[b] + t.size[b]] += 1
    assert (cover == 1).all()
    assert P.adm.any() and (~P.adm).any()


def emulate_matvec(W, pool, v1, v2, n, nb=1):
    """Phases 1, 1b and 2 with the kernels' indexing; v1, v2 interleaved
    (cell*nb + b)."""
    t = np.zeros((max(W["nt"], 1), nb))
    part = np.zeros((max(W["npart"], 1), nb))
    X1, X2 = v1.reshape(n, nb), v2.reshape(n, nb)
    for mem0, nmem, cb, cc, bn, c0, direct, ktot in W["units"]:
        for vt, k, out, lane0 in W["members"][mem0:mem0 + nmem]:
            acc = np.zeros((k, nb))
            for c in range(cb, cb + cc):
                x = X2[c0 + c - bn] if c >= bn else X1[c0 + c]
                acc += np.outer(pool[vt + c * k: vt + c * k + k], x)
            (t if direct else part)[out:out + k] = acc
    for to, k, p0, nu in W["reduce"]:
        t[to:to + k] = sum(part[p0 + q * k: p0 + q * k + k] for q in range(nu))
    y = np.zeros((3, n, nb))
    items, ptr = W["items"], W["tile_ptr"]
    for q, (R0, mt) in enumerate(W["tiles"]):
        lanes = np.arange(mt)
        for it in items[ptr[q]:ptr[q + 1]]:
            kind, base, ld, plane, cnt, xo = it
            w = {0: X1, 1: t, 2: X2}[kind]
            for c in range(cnt):
                p = base + lanes + c * ld
                for o in range(3):
                    y[o, R0:R0 + mt] += np.outer(pool[p + o * plane], w[xo + c])
    return y.reshape(3, n * nb)


@pytest.mark.parametrize("leaf", [32, 64])
def test_layout_and_worklists_reproduce_dense_product(oracle, leaf):
    mesh = undulating_fault(length=40e3, depth=12e3, nx=44, nz=18, seed=4)
    ev, xg = get_eleVec(mesh.nodelst, mesh.elelst)
    n = mesh.n
    t = build_cluster_tree(xg, leaf=leaf)
    P = build_block_partition(t, eta=2.0)
    perm = t.perm
    K = oracle.traction_block(xg, ev, mesh.triangles(), perm, perm, MU, LAM)  # cluster order
    r0, m = t.start[P.rows], t.size[P.rows]
    c0, nn = t.start[P.cols], t.size[P.cols]
    lr_idx = np.where(P.adm)[0]
    ranks, fac = [], {}
    for bi in lr_idx:
        a, b = slice(r0[bi], r0[bi] + m[bi]), slice(c0[bi], c0[bi] + nn[bi])
        S = np.block([[K[0][a, b], K[3][a, b]], [K[1][a, b], K[4][a, b]], [K[2][a, b], K[5][a, b]]])
        U, s, Vt = np.linalg.svd(S, full_matrices=False)
        k = int((s > 1e-9 * s[0]).sum())
        ranks.append(k)
        fac[bi] = (U[:, :k] * s[:k], Vt[:k])
    ranks = np.array(ranks)
    L = plan_layout(r0, m, c0, nn, P.adm, lr_idx, ranks)
    pool = np.zeros(L["total"])
    for bi, base in zip(L["dense_idx"], L["d_base"]):
        bm, bn = m[bi], nn[bi]
        for op in range(6):
            blk = K[op][r0[bi]:r0[bi] + bm, c0[bi]:c0[bi] + bn]
            pool[base + op * bm * bn: base + (op + 1) * bm * bn] = blk.T.ravel()  # column-major
    tiles, tor = leaf_tables(t)
    for bi, ub, vb, k in zip(L["lr_idx"], L["l_ub"], L["l_vb"], L["k"]):
        U, V = fac[bi]
        seg = leaf_segments(tiles, tor, r0[bi], m[bi])
        pool[ub:ub + U.size] = pack_u(U, seg)   # one panel per leaf of the block rows
        assert np.array_equal(unpack_u(pool[ub:ub + U.size], m[bi], k, seg), U)
        pool[vb:vb + V.size] = V.T.ravel()   # V stored transposed: 2n x k row-major
    W = build_worklists(t, n, r0, m, c0, nn, L["dense_idx"], L["d_base"], L["lr_idx"],
                        L["l_ub"], L["l_vb"], L["k"])
    rng = np.random.default_rng(0)
    v1, v2 = rng.standard_normal(n), rng.standard_normal(n)
    y = emulate_matvec(W, pool, v1, v2, n)
    ref = np.stack([K[0] @ v1 + K[3] @ v2, K[1] @ v1 + K[4] @ v2, K[2] @ v1 + K[5] @ v2])
    scale = np.abs(K).max() * np.abs(np.concatenate([v1, v2])).sum()
    assert np.abs(y - ref).max() <= 1e-8 * scale
    assert len(L["lr_idx"]) > 0 and len(L["dense_idx"]) > 0
    assert len(W["reduce"]) > 0 or leaf == 64  # some V's span several units
    # four interleaved right-hand sides
    V1, V2 = rng.standard_normal((n, 4)), rng.standard_normal((n, 4))
    y4 = emulate_matvec(W, pool, V1.ravel(), V2.ravel(), n, nb=4).reshape(3, n, 4)
    for b in range(4):
        rb = np.stack([K[0] @ V1[:, b] + K[3] @ V2[:, b], K[1] @ V1[:, b] + K[4] @ V2[:, b],
                       K[2] @ V1[:, b] + K[5] @ V2[:, b]])
        assert np.abs(y4[:, :, b] - rb).max() <= 1e-8 * scale


def test_block_diagonal_union_reproduces_each_product(oracle):
    """MultiHMatrix-style merge of the work lists of two meshes (pool, t and
    partial offsets, column and row offsets) reproduces each mesh's dense
    product on its own slice of the concatenated vectors."""
    import types
    from pyquake3d_amd import hmatrix as hm
    parts = []
    for nx, nz, seed in ((20, 8, 1), (14, 6, 2)):
        mesh = undulating_fault(length=20e3, depth=8e3, nx=nx, nz=nz, seed=seed)
        ev, xg = get_eleVec(mesh.nodelst, mesh.elelst)
        n = mesh.n
        t = build_cluster_tree(xg, leaf=16)
        P = build_block_partition(t, eta=2.0)
        K = oracle.traction_block(xg, ev, mesh.triangles(), t.perm, t.perm, MU, LAM)
        r0, m = t.start[P.rows], t.size[P.rows]
        c0, nn = t.start[P.cols], t.size[P.cols]
        lr_idx = np.where(P.adm)[0]
        ranks, fac = [], {}
        for bi in lr_idx:
            a, b = slice(r0[bi], r0[bi] + m[bi]), slice(c0[bi], c0[bi] + nn[bi])
            S = np.block([[K[0][a, b], K[3][a, b]], [K[1][a, b], K[4][a, b]],
                          [K[2][a, b], K[5][a, b]]])
            U, s, Vt = np.linalg.svd(S, full_matrices=False)
            k = int((s > 1e-9 * s[0]).sum())
            ranks.append(k)
            fac[bi] = (U[:, :k] * s[:k], Vt[:k])
        L = plan_layout(r0, m, c0, nn, P.adm, lr_idx, np.array(ranks))
        pool = np.zeros(L["total"])
        for bi, base in zip(L["dense_idx"], L["d_base"]):
            bm, bn = m[bi], nn[bi]
            for op in range(6):
                blk = K[op][r0[bi]:r0[bi] + bm, c0[bi]:c0[bi] + bn]
                pool[base + op * bm * bn: base + (op + 1) * bm * bn] = blk.T.ravel()
        tiles, tor = leaf_tables(t)
        for bi, ub, vb, k in zip(L["lr_idx"], L["l_ub"], L["l_vb"], L["k"]):
            U, V = fac[bi]
            pool[ub:ub + U.size] = pack_u(U, leaf_segments(tiles, tor, r0[bi], m[bi]))
            pool[vb:vb + V.size] = V.T.ravel()
        W = build_worklists(t, n, r0, m, c0, nn, L["dense_idx"], L["d_base"], L["lr_idx"],
                            L["l_ub"], L["l_vb"], L["k"])
        h = types.SimpleNamespace(
            n=n, perm=t.perm, dtype=torch.float64, device=torch.device("cpu"), esize=8,
            pool=torch.from_numpy(pool), units=torch.from_numpy(W["units"]),
            members=torch.from_numpy(W["members"]),
            reduce=torch.from_numpy(W["reduce"]), tiles=torch.from_numpy(W["tiles"]),
            items=torch.from_numpy(W["items"]), tile_ptr=torch.from_numpy(W["tile_ptr"]),
            nt=W["nt"], npart=W["npart"])
        parts.append((h, K))
    Mu = hm.MultiHMatrix([h for h, _ in parts])
    Wm = dict(units=Mu.units.numpy(), members=Mu.members.numpy(), reduce=Mu.reduce.numpy(),
              tiles=Mu.tiles.numpy(),
              items=Mu.items.numpy(), tile_ptr=Mu.tile_ptr.numpy(), nt=Mu.nt, npart=Mu.npart)
    rng = np.random.default_rng(3)
    v1, v2 = rng.standard_normal(Mu.n), rng.standard_normal(Mu.n)
    y = emulate_matvec(Wm, Mu.pool.numpy(), v1, v2, Mu.n)
    for b, (h, K) in enumerate(parts):
        o, n = int(Mu.off[b]), h.n
        x1, x2 = v1[o:o + n], v2[o:o + n]
        ref = np.stack([K[0] @ x1 + K[3] @ x2, K[1] @ x1 + K[4] @ x2, K[2] @ x1 + K[5] @ x2])
        scale = np.abs(K).max() * np.abs(np.concatenate([x1, x2])).sum()
        assert np.abs(y[:, o:o + n] - ref).max() <= 1e-8 * scale


def test_leaf_panels_of_u():
    """U of a block spanning several leaves: every leaf's rows of all three
    planes and all ranks form one contiguous panel, in leaf order, plane-major
    inside (each plane a column-major ls x k sub-panel), and a single-leaf
    block is three column-major m x k planes one after another."""
    U = np.arange(3 * 10 * 4, dtype=np.float64).reshape(30, 4)   # m = 10, k = 4
    seg = np.array([[0, 3], [3, 5], [8, 2]])
    flat = pack_u(U, seg)
    for ro, ls in seg:
        panel = flat[3 * 4 * ro:3 * 4 * (ro + ls)].reshape(3, 4, ls)
        for o in range(3):
            assert np.array_equal(panel[o].T, U[o * 10 + ro:o * 10 + ro + ls, :])
    assert np.array_equal(unpack_u(flat, 10, 4, seg), U)
    assert np.array_equal(pack_u(U, np.array([[0, 10]])),
                          U.reshape(3, 10, 4).transpose(0, 2, 1).ravel())


def test_xcd_order_is_a_bijection():
    """Host restatement of the kernels' workgroup -> work-index map: each of
    the 8 XCDs (workgroup b runs on XCD b % 8) gets one contiguous range of
    work indices, and every index in [0, g) is produced once."""
    def xcd_order(b, g):
        q, r = divmod(g, 8)
        x, i = b % 8, b // 8
        return x * (q + 1) + i if x < r else r * (q + 1) + (x - r) * q + i
    for g in (1, 7, 8, 9, 100, 2048, 2053):
        w = np.array([xcd_order(b, g) for b in range(g)])
        assert np.array_equal(np.sort(w), np.arange(g))
        for x in range(min(g, 8)):
            mine = np.sort(w[x::8])
            assert np.array_equal(mine, np.arange(mine[0], mine[0] + len(mine)))


def test_unit_packing_covers_every_block_chunk_once():
    """Blocks sharing source columns share wavefronts: total rank <= 64 for a
    unit of several members, lanes consecutive from 0, every (block, column
    chunk) in exactly one unit, output rows direct for one chunk and partial
    rows u*k apart otherwise; units reading the same columns are adjacent."""
    from pyquake3d_amd.hmatrix import UNIT_COLS, pack_units
    rng = np.random.default_rng(3)
    groups, expect = {}, {}
    toff = poff = 0
    for g in range(40):
        bc0, bn = 1000 * g, int(rng.choice([30, 200, 300, 700]))
        for j in range(int(rng.integers(1, 9))):
            k = int(rng.choice([3, 10, 17, 31, 40, 64, 90]))
            vb = 10 ** 6 * g + 10 ** 4 * j
            nu = -(-2 * bn // UNIT_COLS)
            groups.setdefault((bc0, bn), []).append((vb, k, toff, poff))
            expect[vb] = (k, toff, poff, nu)
            toff += k
            poff += nu * k if nu > 1 else 0
    units, members = pack_units(groups)
    seen = {}
    for mem0, nmem, cb, cc, bn, c0, direct, ktot in units:
        ms = members[mem0:mem0 + nmem]
        assert ktot == ms[:, 1].sum() and (nmem == 1 or ktot <= 64)
        assert np.array_equal(ms[:, 3], np.concatenate([[0], np.cumsum(ms[:-1, 1])]))
        assert cb % UNIT_COLS == 0 and 0 < cc <= UNIT_COLS and cb + cc <= 2 * bn
        for vb, k, out, lane0 in ms:
            kk, to, po, nu = expect[vb]
            u = cb // UNIT_COLS
            assert k == kk and direct == (nu == 1)
            assert out == (to if nu == 1 else po + u * k)
            seen[(vb, u)] = seen.get((vb, u), 0) + 1
    assert seen == {(vb, u): 1 for vb, (k, to, po, nu) in expect.items() for u in range(nu)}
    key = units[:, 5] * 10 ** 6 + units[:, 2]
    assert np.all(np.diff(key) >= 0)
    assert len(units) < sum(nu for (k, to, po, nu) in expect.values())   # packing happened


@pytest.mark.parametrize("nrhs", [1, 2, 4, 8, 16])
def test_reduce_rows_cover_each_reduction_once(nrhs):
    """Phase-1b rows: every rank of every reduction in exactly one row, rows of
    max(1, 64 // nrhs) ranks, reduction fields copied."""
    from pyquake3d_amd.hmatrix import reduce_rows
    base = np.array([[0, 5, 0, 2], [5, 64, 10, 3], [69, 130, 202, 4], [199, 1, 722, 2]])
    rows = reduce_rows(base, nrhs)
    lc = max(1, 64 // nrhs)
    seen = {}
    for to, k, p0, nu, l0, cnt in rows:
        match = [r for r in base if r[0] == to]
        assert len(match) == 1 and tuple(match[0]) == (to, k, p0, nu)
        assert 1 <= cnt <= lc and l0 % lc == 0 and l0 + cnt <= k
        for l in range(l0, l0 + cnt):
            seen[(to, l)] = seen.get((to, l), 0) + 1
    assert seen == {(r[0], l): 1 for r in base for l in range(r[1])}
    assert reduce_rows(np.zeros((0, 4), dtype=np.int64), nrhs).shape == (0, 6)


def test_hmatrix_file_round_trip(tmp_path, oracle):
    """HMatrix.save / HMatrix.load (safetensors, versioned header) keep the
    pool, the work lists, the permutation and the block tables exactly; the
    loaded tables reproduce the product; another format is refused."""
    import types
    from pyquake3d_amd import _lib
    from pyquake3d_amd.hmatrix import HMatrix, HStats
    mesh = undulating_fault(length=20e3, depth=8e3, nx=18, nz=7, seed=5)
    ev, xg = get_eleVec(mesh.nodelst, mesh.elelst)
    n = mesh.n
    t = build_cluster_tree(xg, leaf=16)
    P = build_block_partition(t, eta=2.0)
    K = oracle.traction_block(xg, ev, mesh.triangles(), t.perm, t.perm, MU, LAM)
    r0, m = t.start[P.rows], t.size[P.rows]
    c0, nn = t.start[P.cols], t.size[P.cols]
    lr_idx = np.where(P.adm)[0]
    ranks, fac = [], {}
    for bi in lr_idx:
        a, b = slice(r0[bi], r0[bi] + m[bi]), slice(c0[bi], c0[bi] + nn[bi])
        S = np.block([[K[0][a, b], K[3][a, b]], [K[1][a, b], K[4][a, b]], [K[2][a, b], K[5][a, b]]])
        U, s, Vt = np.linalg.svd(S, full_matrices=False)
        k = int((s > 1e-9 * s[0]).sum())
        ranks.append(k)
        fac[bi] = (U[:, :k] * s[:k], Vt[:k])
    L = plan_layout(r0, m, c0, nn, P.adm, lr_idx, np.array(ranks))
    pool = np.zeros(L["total"])
    for bi, base in zip(L["dense_idx"], L["d_base"]):
        bm, bn = m[bi], nn[bi]
        for op in range(6):
            blk = K[op][r0[bi]:r0[bi] + bm, c0[bi]:c0[bi] + bn]
            pool[base + op * bm * bn: base + (op + 1) * bm * bn] = blk.T.ravel()
    tiles, tor = leaf_tables(t)
    for bi, ub, vb, k in zip(L["lr_idx"], L["l_ub"], L["l_vb"], L["k"]):
        U, V = fac[bi]
        pool[ub:ub + U.size] = pack_u(U, leaf_segments(tiles, tor, r0[bi], m[bi]))
        pool[vb:vb + V.size] = V.T.ravel()
    W = build_worklists(t, n, r0, m, c0, nn, L["dense_idx"], L["d_base"], L["lr_idx"],
                        L["l_ub"], L["l_vb"], L["k"])
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    h = types.SimpleNamespace(
        n=n, nt=W["nt"], npart=W["npart"], esize=8, halfspace=False, mu=MU, lambda_=LAM,
        pool=T(pool), units=T(W["units"]), members=T(W["members"]), reduce=T(W["reduce"]),
        tiles=T(W["tiles"]), tile_ptr=T(W["tile_ptr"]), items=T(W["items"]), perm=t.perm,
        leaf_tiles=tiles, tile_of_row=tor, _BLOCK_KEYS=HMatrix._BLOCK_KEYS,
        FORMAT=HMatrix.FORMAT, VERSION=HMatrix.VERSION,
        blocks=dict(r0=r0, m=m, c0=c0, n=nn, dense_idx=L["dense_idx"], d_base=L["d_base"],
                    lr_idx=L["lr_idx"], ranks=L["k"], l_ub=L["l_ub"], l_vb=L["l_vb"]),
        stats=HStats(n=n, n_dense=len(L["dense_idx"]), n_lowrank=len(L["lr_idx"]),
                     dense_entries=1, lowrank_entries=2, ranks_mean=3.5, ranks_max=9,
                     fallback_dense=0, bytes=0, build_s={}))
    f = tmp_path / "h.safetensors"
    HMatrix.save(h, f, model="test")
    g = HMatrix.load(f, device="cpu")
    for k in ("pool", "units", "members", "reduce", "tiles", "tile_ptr", "items"):
        assert torch.equal(getattr(g, k), getattr(h, k)), k
    assert np.array_equal(g.perm, t.perm) and np.array_equal(g.perm[g.iperm], np.arange(n))
    for k in HMatrix._BLOCK_KEYS:
        assert np.array_equal(g.blocks[k], h.blocks[k]), k
    assert (g.n, g.nt, g.npart, g.esize, g.meta) == (n, W["nt"], W["npart"], 8, {"model": "test"})
    assert g.stats.ranks_mean == 3.5 and g.dtype == torch.float64
    Wg = dict(units=g.units.numpy(), members=g.members.numpy(), reduce=g.reduce.numpy(),
              tiles=g.tiles.numpy(), items=g.items.numpy(), tile_ptr=g.tile_ptr.numpy(),
              nt=g.nt, npart=g.npart)
    rng = np.random.default_rng(2)
    v1, v2 = rng.standard_normal(n), rng.standard_normal(n)
    assert np.array_equal(emulate_matvec(Wg, g.pool.numpy(), v1, v2, n),
                          emulate_matvec(W, pool, v1, v2, n))
    from safetensors.torch import save_file
    save_file({"x": torch.zeros(1)}, str(tmp_path / "other.safetensors"), metadata={"format": "x"})
    with pytest.raises(ValueError):
        HMatrix.load(tmp_path / "other.safetensors", device="cpu")
