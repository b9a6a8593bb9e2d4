"""Host-side bounds check of every device access the H-matrix kernels make.

HMatrix runs its real host pipeline on CPU tensors with the kernel layer
replaced by a checker: for each launch it recomputes, from the same tables the
kernel reads, the extreme index of every buffer access and asserts it lies
inside the buffer (plus a random ACA rank outcome, including non-converged
blocks). Mirrors the index arithmetic of pyquake3d_amd/csrc/hmat_kernels.hip."""
import numpy as np
import pytest
import torch

from pyquake3d_amd import _lib
from pyquake3d_amd import hmatrix as hm
from pyquake3d_amd.mesh import undulating_fault

from cluster_emulation import fake_block_level, fake_tree_level


LC = None


def rows(t, name):
    """Rows of a table as the kernel reads them: flat buffer, the library's
    own row width, no reliance on the tensor's shape."""
    w = LC[name]
    flat = t.reshape(-1).numpy()
    assert flat.size % w == 0, (name, flat.size, w)
    return flat.reshape(-1, w)


class Checker:
    def __init__(self, rng, fail_frac=0.1):
        self.reg = {}
        self.rng = rng
        self.fail_frac = fail_frac
        self.aca_caps = {}
        self.calls = []

    def ptr(self, t):
        if t is None:
            return None
        key = len(self.reg) + 1
        self.reg[key] = t
        return _lib.DevPtr(key)

    def T(self, k):
        return self.reg[k.value]

    def call(self, name, *a):
        # the real argument-kind check of the binding, then the bounds check
        _lib._checked_args(name, getattr(_lib.load(), name), a)
        self.calls.append(name)
        getattr(self, name)(*a)

    def _recs(self, recv, rec):
        N = self.T(recv).numel() // LC["recv"]
        assert self.T(rec).numel() // LC["rec"] == N
        return N

    def _hrecs(self, hrec, N):
        if hrec is not None and hrec != 0:
            assert self.T(hrec).numel() == LC["hrec"] * N

    def pq3d_hmat_dense_assemble(self, tab, nb, recv, rec, hrec, mu, lam, pool, esize, stream):
        tab = rows(self.T(tab), "dense_block")
        N = self._recs(recv, rec)
        self._hrecs(hrec, N)
        assert len(tab) >= nb
        P = self.T(pool).numel()
        for b in range(nb):
            r0, m, c0, n, base = tab[b, :5]
            assert 0 <= r0 and r0 + m <= N and 0 <= c0 and c0 + n <= N and m > 0 and n > 0
            assert base >= 0 and base + 6 * m * n <= P

    def pq3d_tree_level(self, pts, n, perm, tmp, nodes, nnodes, leaf, box, nl, stream):
        # k_tree_level: workgroup q reads nodes[2q..2q+1], perm/tmp[s, s+m),
        # pts[3 perm[i] + k], writes box[6q..6q+5] and nl[q]
        P, pm, tp = self.T(pts), self.T(perm), self.T(tmp)
        nd = self.T(nodes).reshape(-1).numpy()
        assert P.numel() == 3 * n and pm.numel() == n and tp.numel() == n
        assert nd.size >= 2 * nnodes and leaf >= 1
        assert self.T(box).numel() >= 6 * nnodes and self.T(nl).numel() >= nnodes
        s, m = nd[0:2 * nnodes:2], nd[1:2 * nnodes:2]
        assert (s >= 0).all() and (m >= 1).all() and (s + m <= n).all()
        order = np.argsort(s)                     # the nodes' ranges are disjoint
        assert (s[order][1:] >= (s + m)[order][:-1]).all()
        pv = pm.numpy()
        for a, b in zip(s.tolist(), m.tolist()):
            assert pv[a:a + b].min() >= 0 and pv[a:a + b].max() < n
        fake_tree_level(P.view(-1, 3), pm, tp, self.T(nodes).view(-1, 2), leaf,
                        self.T(box).view(-1, 6), self.T(nl))

    def pq3d_block_level(self, pairs, npairs, node, box, eta, min_size, depth, leafs, nxt,
                         counts, stream):
        # k_block_level: thread p reads pairs[3p..3p+2], node[4a..], box[6a..]
        # of both nodes, and writes at most one leaf row (< npairs) or four
        # child rows (< 4 npairs)
        pr = self.T(pairs).reshape(-1, 3).numpy()[:npairs]
        nd = self.T(node).reshape(-1, 4).numpy()
        assert self.T(box).numel() == 6 * len(nd)
        assert ((pr[:, :2] >= 0) & (pr[:, :2] < len(nd))).all()
        assert 0 <= depth < 30
        assert self.T(leafs).numel() >= 4 * npairs and self.T(nxt).numel() >= 12 * npairs
        assert self.T(counts).numel() >= 2
        assert (self.T(counts) == 0).all()
        kids = nd[pr[:, :2]][:, :, 2:]
        split = (kids >= 0).all(axis=(1, 2))
        assert (kids[split] < len(nd)).all()
        fake_block_level(self.T(pairs).view(-1, 3)[:npairs], self.T(node).view(-1, 4),
                         self.T(box).view(-1, 6), eta, min_size, depth,
                         self.T(leafs).view(-1, 4), self.T(nxt).view(-1, 3), self.T(counts))

    def pq3d_hmat_aca(self, tab, wsoff, nb, recv, rec, hrec, mu, lam, eps, kmax, wU, wV, flags,
                      rank, stream):
        tab_a = rows(self.T(tab), "lowrank_block")
        ws = rows(self.T(wsoff), "aca_ws")
        N = self._recs(recv, rec)
        self._hrecs(hrec, N)
        nU, nV, nF = self.T(wU).numel(), self.T(wV).numel(), self.T(flags).numel()
        assert self.T(rank).numel() >= nb
        r = np.empty(nb, dtype=np.int64)
        for b in range(nb):
            r0, m, c0, n = tab_a[b, :4]
            kcap = min(int(tab_a[b, 6]), kmax)
            assert kcap >= 1
            assert r0 + m <= N and c0 + n <= N
            assert ws[b, 0] + kcap * 3 * m <= nU
            assert ws[b, 1] + kcap * 2 * n <= nV
            assert ws[b, 2] + 3 * m + 2 * n <= nF
            r[b] = -1 if self.rng.random() < self.fail_frac else self.rng.integers(1, kcap + 1)
            self.aca_caps[(int(r0), int(c0))] = kcap
        self.T(rank)[:nb] = torch.from_numpy(r)

    def pq3d_hmat_aca_pack(self, tab, wsoff, nb, tiles, tor, wU, wV, pool, esize, stream):
        tab_a = rows(self.T(tab), "lowrank_block")
        ws = rows(self.T(wsoff), "aca_ws")
        tl = rows(self.T(tiles), "tile")
        tor_a = self.T(tor).numpy()
        P = self.T(pool).numel()
        for b in range(nb):
            r0, m, c0, n, ub, vb, k = tab_a[b, :7]
            if k <= 0:
                continue
            assert k <= self.aca_caps[(int(r0), int(c0))]
            assert ws[b, 0] + 3 * m * k <= self.T(wU).numel()
            assert ws[b, 1] + 2 * n * k <= self.T(wV).numel()
            assert 0 <= ub and ub + 3 * m * k <= P and 0 <= vb and vb + 2 * n * k <= P
            # the kernel's U destination, every row of the block: inside the
            # block's 3mk range, and a bijection onto it
            assert 0 <= r0 and r0 + m <= len(tor_a)
            q = tor_a[r0:r0 + m]
            assert (q >= 0).all() and (q < len(tl)).all()
            ro, ls = tl[q, 0] - r0, tl[q, 1]
            i = np.arange(m)
            assert (ro >= 0).all() and (i >= ro).all() and (i < ro + ls).all()
            assert (ro + ls <= m).all()
            dst = (3 * k * ro[None, None, :] + np.arange(3)[None, :, None] * k * ls[None, None, :]
                   + np.arange(k)[:, None, None] * ls[None, None, :] + (i - ro)[None, None, :])
            assert np.array_equal(np.sort(dst.ravel()), np.arange(3 * m * k))

    def pq3d_hmat_matvec(self, units, nun, members, red, nred, tiles, tile_ptr, ntiles, items,
                         pool,
                         esize, nrhs, v1, v2, t, part, y1, y2, y3, ybs, stream):
        P = self.T(pool).numel()
        assert nrhs in (1, 2, 4, 8, 16)
        N = self.T(v1).numel() // nrhs
        assert self.T(v2).numel() == N * nrhs
        assert min(self.T(y).numel() for y in (y1, y2, y3)) >= N * nrhs
        # instance-major outputs: (row, b) at b*ybs + row
        assert ybs == 0 or (ybs >= N and (nrhs - 1) * ybs + N <= self.T(y1).numel())
        nt = self.T(t).numel() // nrhs
        npart = self.T(part).numel() // nrhs
        if nun:
            un = rows(self.T(units), "unit")
            assert len(un) >= nun
            mb = rows(self.T(members), "member")
            for mem0, nmem, cb, cc, bn, c0, direct, ktot in un[:nun]:
                assert nmem >= 1 and 0 <= mem0 and mem0 + nmem <= len(mb)
                assert cc > 0 and 0 <= cb and cb + cc <= 2 * bn
                assert c0 >= 0 and c0 + bn <= N
                ms = mb[mem0:mem0 + nmem]
                assert ktot == ms[:, 1].sum() and (nmem == 1 or ktot <= 64)
                lanes = np.concatenate([[0], np.cumsum(ms[:-1, 1])])
                assert np.array_equal(ms[:, 3], lanes)   # the kernel's lane -> member map
                for vt, k, out, lane0 in ms:
                    assert k > 0
                    assert vt >= 0 and vt + (cb + cc) * k <= P      # V^T rows cb..cb+cc
                    assert out >= 0 and out + k <= (nt if direct else npart)
        if nred:
            rd = rows(self.T(red), "reduce")
            assert len(rd) >= nred
            for to, k, p0, nu, l0, lc in rd[:nred]:
                assert to + k <= nt and p0 >= 0 and p0 + nu * k <= npart
                assert 0 <= l0 and lc >= 1 and l0 + lc <= k
        tl = rows(self.T(tiles), "tile")
        tp = self.T(tile_ptr).numpy()
        it = rows(self.T(items), "item")
        assert len(tl) >= ntiles and len(tp) >= ntiles + 1 and tp[ntiles] <= len(it)
        for q in range(ntiles):
            R0, mt = tl[q]
            assert 0 < mt <= 64 and R0 + mt <= N
            for kind, base, ld, plane, cnt, xo in it[tp[q]:tp[q + 1]]:
                assert cnt > 0 and base >= 0
                hi = base + (cnt - 1) * ld + (mt - 1)
                assert kind in (0, 1, 2) and hi + 2 * plane < P and xo >= 0
                assert xo + cnt <= (nt if kind == 1 else N)


@pytest.mark.parametrize("leaf,eta,dtype,hs", [(32, 2.0, torch.float32, False),
                                               (64, 2.0, torch.float64, False),
                                               (32, 0.0, torch.float64, False),
                                               (32, 2.0, torch.float32, True)])
def test_hmatrix_device_accesses_in_bounds(monkeypatch, leaf, eta, dtype, hs):
    global LC
    LC = _lib.layout_constants()
    chk = Checker(np.random.default_rng(leaf))
    monkeypatch.setattr(_lib, "call", chk.call)
    monkeypatch.setattr(_lib, "ptr", chk.ptr)
    monkeypatch.setattr(_lib, "stream_handle", lambda s=None: None)
    mesh = undulating_fault(length=40e3, depth=12e3, nx=44, nz=18, seed=4)
    H = hm.HMatrix.from_mesh(mesh, 32e9, 32e9, leaf=leaf, eta=eta, eps=1e-4, dtype=dtype,
                             device="cpu", aca_budget_gb=0.002, halfspace=hs)
    x = torch.zeros(mesh.n, dtype=torch.float64)
    H.matvec(x, x)
    x8 = torch.zeros(mesh.n * 8, dtype=torch.float64)
    H.matvec(x8, x8, nrhs=8)
    H.matvec(x8, x8, nrhs=8, instance_major=True)
    assert "pq3d_hmat_matvec" in chk.calls
    if eta > 0:
        assert chk.calls.count("pq3d_hmat_aca") > 1   # the workspace budget forced batches


def test_tde_and_blocktree_call_argument_kinds(monkeypatch):
    """Every Python path into the library passes pointers where the C-ABI
    takes pointers and integers/doubles elsewhere (checked by the binding
    itself), and the traction block's output bound matches the tensor."""
    from pyquake3d_amd.blocktree import createHmatrix
    from pyquake3d_amd.mesh import get_eleVec
    from pyquake3d_amd.tde import DeviceMesh
    seen = []
    reg = {}

    def ptr(t):
        if t is None:
            return None
        reg[len(reg) + 1] = t
        return _lib.DevPtr(len(reg))

    def call(name, *a):
        _lib._checked_args(name, getattr(_lib.load(), name), a)
        seen.append(name)
        if name.startswith("pq3d_tde_traction_block"):
            out = reg[a[10].value]
            m, n, si, sj, plane, out_len = a[4], a[7], a[11], a[12], a[13], a[14]
            assert out_len == out.numel()
            assert 5 * plane + (m - 1) * si + (n - 1) * sj < out_len
        if name == "pq3d_hmat_aca":
            reg[a[13].value][:] = 3

    monkeypatch.setattr(_lib, "call", call)
    monkeypatch.setattr(_lib, "ptr", ptr)
    monkeypatch.setattr(_lib, "stream_handle", lambda s=None: None)
    mesh = undulating_fault(length=12e3, depth=6e3, nx=12, nz=6, seed=1)
    dm = DeviceMesh.from_mesh(mesh, 32e9, 32e9, device="cpu")
    dm.traction_block()
    dm.traction_block([0, 3, 5], [1, 2])
    dm.traction_block(None, None, row0=2, m=7, col0=3, n=9, order="ji")
    dm.traction_block([1, 2], [3], dtype=torch.float32)
    ev, xg = get_eleVec(mesh.nodelst, mesh.elelst)
    tree = createHmatrix(xg, mesh.nodelst, mesh.elelst, ev, 32e9, 32e9, False, mini_leaf=8,
                         device="cpu")
    tree.dm.traction_block = lambda r, c: torch.zeros((6, len(r), len(c)), dtype=torch.float64)
    for b in tree.blocks_to_process:
        tree.ACA_worker(b)
    assert "pq3d_hmat_aca" in seen and seen.count("pq3d_tde_traction_block") == 3
