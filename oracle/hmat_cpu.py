"""CPU baseline of the H-matrix time step -- TEST INFRASTRUCTURE ONLY.

Only bench.py's CPU-baseline leg uses it. It restates the reference's CPU
path for one step: ``blocks_process_MVM`` (src/Hmatrix.py:1445-1538), a Python
loop over leaf blocks with numpy FP64 products (dense ``Mf @ x`` or
``U @ (V @ x)``) scattered into y, over the same blocks and factors the GPU
build produced (copied to host FP64), followed by the oracle RK45 step
(oracle/qdsim_oracle.py).

``ScatterMatvec`` runs the same product on several host processes, as the
reference's MPI ranks do.

Pinned by tests/test_oracle_hmat_cpu.py: on pools laid out as the GPU build
lays them out, an all-dense partition equals the dense FP64 oracle product
to 1e-12, SVD-truncated far-field blocks stay within their truncation, the
two-plane product gives the first two rows, and the scattered product on
several processes equals the single-process one to rounding.
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


def _scatter_worker(conn, dense, lowrank, xs, ys):
    """One rank of the block scatter: its blocks' contributions into its own
    partial output, on request."""
    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=1):
        while True:
            planes = conn.recv()
            if planes is None:
                break
            v1, v2 = xs[0], xs[1]
            ys[:] = 0.0
            for r0, m, c0, n, M in dense:
                x = np.concatenate([v1[c0:c0 + n], v2[c0:c0 + n]])
                ys[:planes, r0:r0 + m] += (M[:planes * m] @ x).reshape(planes, m)
            for r0, m, c0, n, U, V in lowrank:
                x = np.concatenate([v1[c0:c0 + n], v2[c0:c0 + n]])
                ys[:planes, r0:r0 + m] += (U[:planes * m] @ (V @ x)).reshape(planes, m)
            conn.send(1)


class ScatterMatvec:
    """The reference's MPI layout of the product on ``nproc`` host processes:
    blocks are dealt to ranks (src/Hmatrix.py:1376-1441; here by greedy
    balance of stored entries), every rank applies its blocks
    (blocks_process_MVM, src/Hmatrix.py:1445-1538) into a partial vector,
    and the partials are summed as the reduce of src/QDsim.py:920-922 does.
    Workers are forked copies of this process (no exec) holding their
    blocks; x and the partials live in shared memory. Each worker runs its
    numpy with one BLAS thread. Call ``close()`` when done."""

    def __init__(self, C: CpuHMatrix, nproc: int):
        import multiprocessing as mp
        ctx = mp.get_context("fork")
        self.n, self.nproc = C.n, int(nproc)
        cost = [M.size for *_, M in C.dense] + [U.size + V.size for *_, U, V in C.lowrank]
        items = [("d", i) for i in range(len(C.dense))] + [("l", i) for i in range(len(C.lowrank))]
        load = np.zeros(self.nproc)
        mine = [([], []) for _ in range(self.nproc)]
        for k in np.argsort(cost)[::-1]:
            p = int(np.argmin(load))
            load[p] += cost[k]
            kind, i = items[k]
            (mine[p][0].append(C.dense[i]) if kind == "d" else mine[p][1].append(C.lowrank[i]))
        self._xs = np.frombuffer(ctx.RawArray("d", 2 * self.n), dtype=np.float64).reshape(2, self.n)
        self._ys = np.frombuffer(ctx.RawArray("d", self.nproc * 3 * self.n),
                                 dtype=np.float64).reshape(self.nproc, 3, self.n)
        self._conns, self._procs = [], []
        for p in range(self.nproc):
            a, b = ctx.Pipe()
            pr = ctx.Process(target=_scatter_worker,
                             args=(b, mine[p][0], mine[p][1], self._xs, self._ys[p]), daemon=True)
            pr.start()
            self._conns.append(a)
            self._procs.append(pr)

    def matvec(self, v1, v2, planes=3):
        self._xs[0] = v1
        self._xs[1] = v2
        for c in self._conns:
            c.send(int(planes))
        for c in self._conns:
            c.recv()
        return self._ys.sum(axis=0)

    def close(self):
        for c in self._conns:
            try:
                c.send(None)
            except (BrokenPipeError, OSError):
                pass
        for p in self._procs:
            p.join(timeout=10)
            if p.is_alive():
                p.terminate()
        self._conns, self._procs = [], []
