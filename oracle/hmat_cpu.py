"""CPU baseline of the H-matrix time step -- TEST INFRASTRUCTURE ONLY.

Only bench.py's CPU-baseline leg uses it. It restates the reference's CPU
path for one step: ``blocks_process_MVM`` (src/Hmatrix.py:1445-1538), a Python
loop over leaf blocks with numpy FP64 products (dense ``Mf @ x`` or
``U @ (V @ x)``) scattered into y, over the same blocks and factors the GPU
build produced (copied to host FP64), followed by the oracle RK45 step
(oracle/qdsim_oracle.py).

Pinned by tests/test_oracle_hmat_cpu.py: on pools laid out as the GPU build
lays them out, an all-dense partition equals the dense FP64 oracle product
to 1e-12, SVD-truncated far-field blocks stay within their truncation, and
the two-plane product gives the first two rows.
"""
from __future__ import annotations

import numpy as np


class CpuHMatrix:
    def __init__(self, H):
        B = H.blocks
        pool = H.pool.double().cpu().numpy()
        self.n = H.n
        self.dense = []
        for bi, base in zip(B["dense_idx"], B["d_base"]):
            r0, m, c0, n = int(B["r0"][bi]), int(B["m"][bi]), int(B["c0"][bi]), int(B["n"][bi])
            planes = pool[base:base + 6 * m * n].reshape(6, n, m).transpose(0, 2, 1)
            # stacked [[A1s, A1d], [A2s, A2d], [Bs, Bd]] (3m x 2n), FP64, C order
            M = np.block([[planes[0], planes[3]], [planes[1], planes[4]], [planes[2], planes[5]]])
            self.dense.append((r0, m, c0, n, np.ascontiguousarray(M)))
        self.lowrank = []
        for q, bi in enumerate(B["lr_idx"]):
            U, V = H.lowrank_factors(q)
            r0, m, c0, n = int(B["r0"][bi]), int(B["m"][bi]), int(B["c0"][bi]), int(B["n"][bi])
            self.lowrank.append((r0, m, c0, n, np.ascontiguousarray(U), np.ascontiguousarray(V)))

    def matvec(self, v1, v2, planes=3):
        """(AdotV1, AdotV2, dsigma) in cluster order, FP64. ``planes=2``
        applies only the four A operators, as the reference does with a
        fixed normal traction (src/QDsim.py:906-909); dsigma is then zero."""
        y = np.zeros((3, self.n))
        for r0, m, c0, n, M in self.dense:
            x = np.concatenate([v1[c0:c0 + n], v2[c0:c0 + n]])
            y[:planes, r0:r0 + m] += (M[:planes * m] @ x).reshape(planes, m)
        for r0, m, c0, n, U, V in self.lowrank:
            x = np.concatenate([v1[c0:c0 + n], v2[c0:c0 + n]])
            y[:planes, r0:r0 + m] += (U[:planes * m] @ (V @ x)).reshape(planes, m)
        return y
