"""V panels of the matrix-core product (hmatrix.vpanel_plan), on the CPU:
the packing rows move every V^T element of every block to its panel chunk,
and phase 1 over the panel units (partial rows summed by the reduce rows)
gives t = V^T x of every block at its panel t rows."""
import numpy as np
import pytest

from pyquake3d_amd.hmatrix import vpanel_plan


def _random_blocks(rng, ncl=9, nb=40):
    """Low-rank blocks over a few source clusters (c0, n) with ranks 1..90
    and their V^T (2n x k, row-major) at random bases of a pool."""
    sizes = rng.integers(3, 40, ncl)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    cl = rng.integers(0, ncl, nb)
    c0, nn = starts[cl].astype(np.int64), sizes[cl].astype(np.int64)
    k = rng.integers(1, 91, nb).astype(np.int64)
    lr = np.arange(nb, dtype=np.int64)
    size = 2 * nn * k
    gaps = rng.integers(0, 7, nb)
    vb = np.concatenate([[0], np.cumsum(size + gaps)[:-1]]).astype(np.int64)
    pool = rng.standard_normal(int(vb[-1] + size[-1]))
    return c0, nn, lr, k, vb, pool, int(sizes.sum())


def _pack(plan, pool):
    """The packing kernel's addressing, element by element."""
    vp = np.full(plan["size"], np.nan)
    for src, k, n, base, K, koff in plan["tab"]:
        C = 2 * n
        for p in range(C * k):
            c, l = divmod(p, k)
            r = koff + l
            j = r // 64
            kc = min(64, K - 64 * j)
            vp[base + C * 64 * j + c * kc + (r - 64 * j)] = pool[src + p]
    return vp


def _phase1(plan, vp, x1, x2, nrows):
    """Phase 1 as the kernel reads a unit: rows of the chunk (stacked
    columns cb .. cb+cc) times the x1 rows (stacked column < n) or x2 rows."""
    t = np.zeros((plan["nt"], x1.shape[1]))
    part = np.zeros((max(plan["npart"], 1), x1.shape[1]))
    for vb, kc, out, cb, cc, n, c0, direct in plan["units"]:
        A = vp[vb: vb + cc * kc].reshape(cc, kc)
        g = cb + np.arange(cc)
        X = np.where((g < n)[:, None], x1[np.minimum(c0 + g, nrows - 1)],
                     x2[np.clip(c0 + g - n, 0, nrows - 1)])
        (t if direct else part)[out: out + kc] = A.T @ X
    for to, kc, p0, nu in plan["reduce"]:
        t[to: to + kc] = sum(part[p0 + u * kc: p0 + (u + 1) * kc] for u in range(nu))
    return t


@pytest.mark.parametrize("unit_cols", [8, 64, 2048])
def test_panel_phase1_equals_block_products(unit_cols):
    rng = np.random.default_rng(11)
    c0, nn, lr, k, vb, pool, nrows = _random_blocks(rng)
    plan = vpanel_plan(c0, nn, lr, k, vb, unit_cols=unit_cols)
    assert plan["nt"] == k.sum()
    vp = _pack(plan, pool)
    # every panel element written once; panels start on 128-byte boundaries,
    # the padding between them is never written
    assert int((~np.isnan(vp)).sum()) == int((2 * nn * k).sum()) <= plan["size"]
    assert np.all(plan["tab"][:, 3] % 32 == 0)
    x1, x2 = rng.standard_normal((nrows, 3)), rng.standard_normal((nrows, 3))
    t = _phase1(plan, vp, x1, x2, nrows)
    for b in range(len(lr)):
        Vt = pool[vb[b]: vb[b] + 2 * nn[b] * k[b]].reshape(2 * nn[b], k[b])
        x = np.concatenate([x1[c0[b]: c0[b] + nn[b]], x2[c0[b]: c0[b] + nn[b]]])
        np.testing.assert_allclose(t[plan["toff"][b]: plan["toff"][b] + k[b]], Vt.T @ x,
                                   rtol=1e-12, atol=1e-12)
    # units respect the limits the kernel assumes
    U = plan["units"]
    assert np.all((U[:, 1] >= 1) & (U[:, 1] <= 64) & (U[:, 4] >= 1) & (U[:, 4] <= unit_cols))


def test_panel_t_rows_are_a_permutation_of_block_ranks():
    rng = np.random.default_rng(3)
    c0, nn, lr, k, vb, pool, _ = _random_blocks(rng, ncl=4, nb=25)
    plan = vpanel_plan(c0, nn, lr, k, vb)
    rows = np.concatenate([np.arange(plan["toff"][b], plan["toff"][b] + k[b])
                           for b in range(len(lr))])
    assert np.array_equal(np.sort(rows), np.arange(k.sum()))
    # blocks of one source cluster are side by side, in block order
    for key in set(zip(c0.tolist(), nn.tolist())):
        mem = [b for b in range(len(lr)) if (c0[b], nn[b]) == key]
        offs = plan["toff"][mem]
        assert np.array_equal(offs, offs[0] + np.concatenate([[0], np.cumsum(k[mem])[:-1]]))

