"""Step stream of the 16-RHS matrix-core product (hmatrix.mfs_steps), on the
CPU: a numpy product over the step records, read the way the kernel reads
them (4-column steps, masked columns, wavefront ranges), equals the product
over the work items it was cut from, for every tile order and wavefront
split; record fields and limits."""
import numpy as np
import pytest

from pyquake3d_amd.hmatrix import STEP_CNT_SHIFT, mfs_steps


def _random_worklist(rng, ntiles=7, n=300, nt=90):
    """Tiles of 4..64 rows over [0, n), items of kinds 0/1/2 with random
    strides, column counts 1..32 and weight rows inside x1, x2 or t."""
    rows = rng.integers(4, 65, ntiles)
    starts = np.concatenate([[0], np.cumsum(rows)[:-1]])
    tiles = np.stack([starts, rows], axis=1).astype(np.int64)
    items, ptr, base = [], [0], 0
    for q in range(ntiles):
        for _ in range(int(rng.integers(1, 6))):
            kind = int(rng.integers(0, 3))
            cnt = int(rng.integers(1, 33))
            ld = int(rows[q] + rng.integers(0, 9))
            plane = ld * cnt + int(rng.integers(0, 5))
            xoff = int(rng.integers(0, (nt if kind == 1 else n) - cnt + 1))
            items.append((kind, base, ld, plane, cnt, xoff))
            base += 3 * plane + ld
        ptr.append(len(items))
    return (np.array(items, dtype=np.int64), np.array(ptr, dtype=np.int64), tiles, n, nt,
            base)


def _product_items(items, ptr, tiles, pool, W, n, NP=2):
    """y[o, row] of every tile row from the items: sum over the item's
    columns j of pool[base + o plane + j ld + i] * W[row(kind) + xoff + j]."""
    y = np.zeros((NP, int(tiles[:, 0].max() + tiles[:, 1].max()), W.shape[1]))
    rb = {0: 0, 2: n, 1: 2 * n}
    for q, (r0, mt) in enumerate(tiles):
        for kind, base, ld, plane, cnt, xoff in items[ptr[q]:ptr[q + 1]]:
            for o in range(NP):
                A = np.stack([pool[base + o * plane + j * ld: base + o * plane + j * ld + mt]
                              for j in range(cnt)], axis=1)
                y[o, r0:r0 + mt] += A @ W[rb[kind] + xoff: rb[kind] + xoff + cnt]
    return y


def _product_steps(steps, wptr, tiles, waves, pool, W, zrow, NP=2):
    """The kernel's reading of the step stream: per tile and wavefront, each
    step adds its four columns (a column past the step's count reads the
    step's last column with the zero weight row)."""
    y = np.zeros((NP, int(tiles[:, 0].max() + tiles[:, 1].max()), W.shape[1]))
    rec = steps.view(np.uint32)
    for q, (r0, mt) in enumerate(tiles):
        for g in range(waves):
            for s in range(wptr[q * waves + g], wptr[q * waves + g + 1]):
                b, ld, plane, w = (int(x) for x in rec[s])
                cnt = (w >> STEP_CNT_SHIFT) + 1
                row = w & ((1 << STEP_CNT_SHIFT) - 1)
                for kl in range(4):
                    c = min(kl, cnt - 1)
                    wr = row + c if kl < cnt else zrow
                    for o in range(NP):
                        a = pool[b + c * ld + o * plane: b + c * ld + o * plane + mt]
                        y[o, r0:r0 + mt] += np.outer(a, W[wr])
    return y


@pytest.mark.parametrize("order", ["row", "cost", "rowbal"])
@pytest.mark.parametrize("waves", [1, 2, 4])
def test_step_stream_equals_item_product(order, waves):
    rng = np.random.default_rng(5)
    items, ptr, tiles, n, nt, size = _random_worklist(rng)
    # the work lists' tile order is a permutation of the row order
    perm = rng.permutation(len(tiles))
    tiles_p = tiles[perm]
    items_p = np.concatenate([items[ptr[q]:ptr[q + 1]] for q in perm])
    ptr_p = np.concatenate([[0], np.cumsum(np.diff(ptr)[perm])])
    pool = rng.standard_normal(size + 64)
    W = rng.standard_normal((2 * n + nt + 1, 3))
    W[-1] = 0.0
    ref = _product_items(items_p, ptr_p, tiles_p, pool, W, n)
    steps, wptr, tl = mfs_steps(items_p, ptr_p, tiles_p, n, nt, waves=waves, order=order)
    assert steps.dtype == np.int32 and steps.shape[1] == 4
    assert len(wptr) == len(tl) * waves + 1 and wptr[0] == 0 and wptr[-1] == len(steps)
    assert np.all(np.diff(wptr) >= 0)
    if order == "row":
        assert np.array_equal(tl, tiles)
    elif order == "cost":
        assert np.array_equal(tl, tiles_p)
    else:
        # leaf order in eight ranges, padded with empty tiles to a multiple of 8
        assert len(tl) % 8 == 0 and len(tl) >= len(tiles)
        real = tl[tl[:, 1] > 0]
        assert np.array_equal(real[np.argsort(real[:, 0])], tiles)
        # padding tiles own no steps
        per_tile = np.diff(wptr[::waves])
        assert np.all(per_tile[tl[:, 1] == 0] == 0)
    # one step per 4 columns of every item
    assert len(steps) == int(((items[:, 4] + 3) // 4).sum())
    got = _product_steps(steps, wptr, tl, waves, pool, W, 2 * n + nt)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_step_ranges_split_each_tile_evenly():
    rng = np.random.default_rng(1)
    items, ptr, tiles, n, nt, _ = _random_worklist(rng, ntiles=5)
    for waves in (2, 4):
        steps, wptr, _ = mfs_steps(items, ptr, tiles, n, nt, waves=waves, order="cost")
        per = np.diff(wptr).reshape(len(tiles), waves)
        assert np.all(per.max(axis=1) - per.min(axis=1) <= 1)


def test_step_record_limits():
    rng = np.random.default_rng(2)
    items, ptr, tiles, n, nt, _ = _random_worklist(rng, ntiles=3)
    with pytest.raises(ValueError):
        mfs_steps(items, ptr, tiles, n, nt, waves=3)
    with pytest.raises(ValueError):
        mfs_steps(items, ptr, tiles, n, nt, order="random")
    big = items.copy()
    big[0, 1] = 2 ** 32           # pool element beyond the 32-bit record field
    with pytest.raises(ValueError):
        mfs_steps(big, ptr, tiles, n, nt)
    bad = items.copy()
    k = int(np.where(bad[:, 0] == 1)[0][0]) if np.any(bad[:, 0] == 1) else 0
    bad[k, 0], bad[k, 5] = 1, nt   # weight rows past the t rows
    with pytest.raises(ValueError):
        mfs_steps(bad, ptr, tiles, n, nt)
    with pytest.raises(ValueError):
        mfs_steps(items, ptr[:-1], tiles, n, nt)


def _xcd_order(b, g):
    """The kernels' launch-slot -> work-index map (consecutive slots go to
    the eight XCDs in turn; XCD x owns one contiguous range of indices)."""
    q, r = divmod(g, 8)
    x, i = b % 8, b // 8
    return x * (q + 1) + i if x < r else r * (q + 1) + (x - r) * q + i


def test_rowbal_gives_every_xcd_one_run_of_leaves():
    rng = np.random.default_rng(9)
    items, ptr, tiles, n, nt, _ = _random_worklist(rng, ntiles=37)
    _, _, tl = mfs_steps(items, ptr, tiles, n, nt, waves=1, order="rowbal")
    g = len(tl)
    assert sorted(_xcd_order(b, g) for b in range(g)) == list(range(g))
    start_rank = {int(r0): k for k, r0 in enumerate(tiles[:, 0])}
    runs = []
    for x in range(8):
        ranks = [start_rank[int(tl[_xcd_order(b, g), 0])] for b in range(x, g, 8)
                 if tl[_xcd_order(b, g), 1] > 0]
        if ranks:
            assert ranks == list(range(ranks[0], ranks[0] + len(ranks)))
        runs.append(ranks)
    assert sorted(r for run in runs for r in run) == list(range(len(tiles)))
