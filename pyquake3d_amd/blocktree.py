"""Drop-in mirror of the reference's H-matrix Python surface (src/Hmatrix.py).

Same names, argument meanings and return shapes as the reference, backed by
the HIP library:

  createHmatrix(xg, nodelst, elelst, eleVec, mu_, lambda_, halfspace_jud,
                mini_leaf=32, plotHmatrix=False)          src/Hmatrix.py:1676
  Block                                                   src/Hmatrix.py:194
  BlockTree.GetTtstress(Stress, col_cluster)              src/Hmatrix.py:295
  BlockTree.calc_stressfunc_fromC_rowS/rowD(rows, cols)   src/Hmatrix.py:318, :372
  BlockTree.calc_stressfunc_fromCS/CD(rows, cols)         src/Hmatrix.py:425, :472
  BlockTree.calc_rowsS/D, calc_colsS/D                    src/Hmatrix.py:556-578
  BlockTree.ACA_worker(block)                             src/Hmatrix.py:1014
  BlockTree.blocks_process_MVM(xvector, blocks, type)     src/Hmatrix.py:1445

Evaluators run one GPU launch per call (all rows x cols pairs at once) instead
of one ctypes call per row or column. ACA_worker runs the stacked-block ACA of
the library (relative Frobenius stop, eps 1e-4 by default) and returns the
reference's dictionary layout (U_ACA_A1s (m x k), V_ACA_A1s (k x n), ...),
where the strike (S) and dip (D) factors share rows of one stacked
factorisation. blocks_process_MVM accepts either a list of processed Blocks
(numpy FP64 products, as the reference) or an ``HMatrix`` (device product).
``halfspace_jud`` selects the half-space Green's functions (free surface at
z = 0), as the reference's flag does, for the evaluators and ACA_worker.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .cluster import build_block_partition, build_cluster_tree
from .hmatrix import HMatrix
from .tde import DeviceMesh

OPS = ("A1s", "A2s", "Bs", "A1d", "A2d", "Bd")


class Block:
    """Leaf or parent block of the block tree (src/Hmatrix.py:194-264)."""

    def __init__(self, row_cluster, col_cluster, row_index=None, col_index=None, children=None,
                 level=0):
        self.row_cluster = np.asarray(row_cluster)
        self.col_cluster = np.asarray(col_cluster)
        self.row_index = row_index
        self.col_index = col_index
        self.children = children if children is not None else []
        self.level = level
        self.ACA_dictS = {}
        self.ACA_dictD = {}
        self.Mf_A1s = self.Mf_A2s = self.Mf_Bs = []
        self.Mf_A1d = self.Mf_A2d = self.Mf_Bd = []
        self.judproc = False
        self.judsvd = False
        self.judaca = False

    def is_leaf(self):
        return len(self.children) == 0


class BlockTree:
    def __init__(self, root_block, nodelst, elelst, eleVec, mu_, lambda_, xg, halfspace_jud,
                 mini_leaf, device="cuda"):
        self.root_block = root_block
        self.nodelst = np.asarray(nodelst, dtype=np.float64)
        self.elelst = np.asarray(elelst, dtype=np.int64)
        self.eleVec = np.asarray(eleVec, dtype=np.float64)
        self.xg = np.asarray(xg, dtype=np.float64)
        self.mu_ = float(mu_)
        self.lambda_ = float(lambda_)
        self.halfspace_jud = halfspace_jud
        self.mini_leaf = mini_leaf
        self.thread_svd = mini_leaf
        self.aca_eps = 1e-4
        self.device = torch.device(device)
        self.dm = DeviceMesh(self.nodelst, self.elelst, self.mu_, self.lambda_, device=self.device,
                             halfspace=bool(halfspace_jud))
        # receivers use the caller's frames and centroids (as the reference does)
        rv = np.concatenate([self.xg, self.eleVec], axis=1)
        self.dm.recv = torch.from_numpy(np.ascontiguousarray(rv)).to(self.device)
        self.blocks_to_process = []

    # ---------------------------------------------------------- projection
    def GetTtstress(self, Stress, col_cluster):
        """Traction (Tt1, Tt2, Tn) on the frames of elements ``col_cluster``
        from stress tensors Stress (6, k) (src/Hmatrix.py:295-313)."""
        S = np.asarray(Stress, dtype=np.float64)
        ev = self.eleVec[np.asarray(col_cluster)]
        Tr1 = S[0] * ev[:, 6] + S[3] * ev[:, 7] + S[4] * ev[:, 8]
        Tr2 = S[3] * ev[:, 6] + S[1] * ev[:, 7] + S[5] * ev[:, 8]
        Trn = S[4] * ev[:, 6] + S[5] * ev[:, 7] + S[2] * ev[:, 8]
        return np.stack([Tr1 * ev[:, 0] + Tr2 * ev[:, 1] + Trn * ev[:, 2],
                         Tr1 * ev[:, 3] + Tr2 * ev[:, 4] + Trn * ev[:, 5],
                         Tr1 * ev[:, 6] + Tr2 * ev[:, 7] + Trn * ev[:, 8]], axis=1)

    # ---------------------------------------------------------- evaluators
    def _six(self, row_cluster, col_cluster):
        return self.dm.traction_block(np.asarray(row_cluster, dtype=np.int64),
                                      np.asarray(col_cluster, dtype=np.int64)).cpu().numpy()

    def calc_stressfunc_fromC_rowS(self, row_cluster, col_cluster):
        K = self._six(row_cluster, col_cluster)
        return K[0], K[1], K[2]

    def calc_stressfunc_fromC_rowD(self, row_cluster, col_cluster):
        K = self._six(row_cluster, col_cluster)
        return K[3], K[4], K[5]

    calc_stressfunc_fromCS = calc_stressfunc_fromC_rowS
    calc_stressfunc_fromCD = calc_stressfunc_fromC_rowD

    def calc_rowsS(self, Istart, Iend, row_cluster, col_cluster):
        return self.calc_stressfunc_fromC_rowS(np.asarray(row_cluster)[Istart:Iend], col_cluster)

    def calc_rowsD(self, Istart, Iend, row_cluster, col_cluster):
        return self.calc_stressfunc_fromC_rowD(np.asarray(row_cluster)[Istart:Iend], col_cluster)

    def calc_colsS(self, Jstart, Jend, row_cluster, col_cluster):
        return self.calc_stressfunc_fromCS(row_cluster, np.asarray(col_cluster)[Jstart:Jend])

    def calc_colsD(self, Jstart, Jend, row_cluster, col_cluster):
        return self.calc_stressfunc_fromCD(row_cluster, np.asarray(col_cluster)[Jstart:Jend])

    # ---------------------------------------------------------------- ACA
    def _aca_block(self, rows, cols, kmax=128):
        """Stacked ACA of one block on gathered rows/cols: (U (3m x k), V (k x 2n)) or None."""
        dev = self.device
        rows = np.asarray(rows, dtype=np.int64)
        cols = np.asarray(cols, dtype=np.int64)
        m, n = len(rows), len(cols)
        recv = self.dm.recv[torch.from_numpy(rows).to(dev)].contiguous()
        rec = self.dm.rec[torch.from_numpy(cols).to(dev)].contiguous()
        hrec = (self.dm.hrec[torch.from_numpy(cols).to(dev)].contiguous() if self.dm.halfspace
                else None)
        kcap = int(min(kmax, 3 * m, 2 * n))
        tab = torch.tensor([[0, m, 0, n, 0, 0, kcap, 0]], dtype=torch.int64, device=dev)
        wsoff = torch.zeros((1, 4), dtype=torch.int64, device=dev)
        wU = torch.empty(kcap * 3 * m, dtype=torch.float64, device=dev)
        wV = torch.empty(kcap * 2 * n, dtype=torch.float64, device=dev)
        flags = torch.empty(3 * m + 2 * n, dtype=torch.uint8, device=dev)
        rank = torch.empty(1, dtype=torch.int64, device=dev)
        _lib.check_width("lowrank_block", tab)
        _lib.check_width("aca_ws", wsoff)
        _lib.check_width("recv", recv)
        _lib.check_width("rec", rec)
        _lib.call("pq3d_hmat_aca", _lib.ptr(tab), _lib.ptr(wsoff), 1, _lib.ptr(recv),
                  _lib.ptr(rec), _lib.ptr(hrec), self.mu_, self.lambda_, float(self.aca_eps), kcap, _lib.ptr(wU),
                  _lib.ptr(wV), _lib.ptr(flags), _lib.ptr(rank), _lib.stream_handle())
        k = int(rank.item())
        if k < 0:
            return None
        U = wU[:3 * m * k].cpu().numpy().reshape(k, 3 * m).T
        V = wV[:2 * n * k].cpu().numpy().reshape(k, 2 * n)
        return U, V

    def ACA_worker(self, block):
        """Compress (both clusters > thread_svd) or evaluate densely one leaf
        block (src/Hmatrix.py:1014-1066)."""
        rows, cols = block.row_cluster, block.col_cluster
        m, n = len(rows), len(cols)
        fac = None
        if m > self.thread_svd and n > self.thread_svd:
            fac = self._aca_block(rows, cols)
        if fac is not None:
            U, V = fac
            block.judaca = True
            block.ACA_dictS = {"U_ACA_A1s": U[:m], "V_ACA_A1s": V[:, :n],
                               "U_ACA_A2s": U[m:2 * m], "V_ACA_A2s": V[:, :n],
                               "U_ACA_Bs": U[2 * m:], "V_ACA_Bs": V[:, :n]}
            block.ACA_dictD = {"U_ACA_A1d": U[:m], "V_ACA_A1d": V[:, n:],
                               "U_ACA_A2d": U[m:2 * m], "V_ACA_A2d": V[:, n:],
                               "U_ACA_Bd": U[2 * m:], "V_ACA_Bd": V[:, n:]}
        else:
            K = self._six(rows, cols)
            (block.Mf_A1s, block.Mf_A2s, block.Mf_Bs,
             block.Mf_A1d, block.Mf_A2d, block.Mf_Bd) = K
        block.judproc = True
        return block

    # -------------------------------------------------------------- matvec
    def blocks_process_MVM(self, xvector, blocks_to_process, type):
        """y = op_type x summed over blocks (src/Hmatrix.py:1445-1538).

        ``blocks_to_process`` is a list of processed Blocks (FP64 numpy
        products, like the reference) or an ``HMatrix`` (device matvec of the
        one operator: the other slip input is zero)."""
        if type not in OPS:
            raise ValueError(f"type must be one of {OPS}")
        x = np.asarray(xvector, dtype=np.float64)
        if isinstance(blocks_to_process, HMatrix):
            H = blocks_to_process
            xc = torch.from_numpy(x[H.perm]).to(H.device)
            z = torch.zeros_like(xc)
            strike = type.endswith("s")
            out = H.matvec(xc if strike else z, z if strike else xc)
            c = {"A1": 0, "A2": 1, "B": 2}[type[:-1]]
            y = np.empty(H.n)
            y[H.perm] = out[c].cpu().numpy()
            return y
        y = np.zeros(len(x))
        comp = type[:-1]
        for b in blocks_to_process:
            x_ = x[b.col_cluster]
            if b.judaca:
                d = b.ACA_dictS if type.endswith("s") else b.ACA_dictD
                Ax = d[f"U_ACA_{type}"] @ (d[f"V_ACA_{type}"] @ x_)
            else:
                Ax = getattr(b, f"Mf_{type}") @ x_
            y[b.row_cluster] += Ax
        return y


def createHmatrix(xg, nodelst, elelst, eleVec, mu_, lambda_, halfspace_jud, mini_leaf=32,
                  plotHmatrix=False, eta=2.0, device="cuda"):
    """Block tree with the reference's rules (src/Hmatrix.py:1676-1714): the
    returned BlockTree's ``blocks_to_process`` holds the unprocessed leaf
    Blocks (row/col clusters in original element numbering)."""
    t = build_cluster_tree(np.asarray(xg, dtype=np.float64), leaf=mini_leaf)
    P = build_block_partition(t, eta=eta, min_size=mini_leaf)
    root = Block(t.perm, t.perm)
    tree = BlockTree(root, nodelst, elelst, eleVec, mu_, lambda_, xg, halfspace_jud, mini_leaf,
                     device=device)
    leaves = []
    for a, b, adm in zip(P.rows, P.cols, P.adm):
        rows = t.perm[t.start[a]:t.start[a] + t.size[a]]
        cols = t.perm[t.start[b]:t.start[b] + t.size[b]]
        leaves.append(Block(rows, cols))
    root.children = leaves
    tree.blocks_to_process = leaves
    return tree
