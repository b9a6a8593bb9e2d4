"""ctypes wrapper of liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the CPU-baseline leg of bench.py use this
module, as the checker. It restates the reference TDE kernel in plain C
(oracle/tde_oracle.c) and exposes it to numpy.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
D = ctypes.POINTER(ctypes.c_double)
L = ctypes.POINTER(ctypes.c_int64)


def _p(a):
    return a.ctypes.data_as(D)


class Oracle:
    def __init__(self, path: Path | str | None = None):
        path = Path(path) if path else HERE / "liboracle.so"
        if not path.exists():
            subprocess.run(["make", "-C", str(HERE), str(path)], check=True,
                           stdout=subprocess.DEVNULL)
        lib = ctypes.CDLL(str(path))
        lib.oracle_TDstressFS_C.argtypes = [D, D, D, ctypes.c_size_t, D, D, D] + [
            ctypes.c_double] * 5 + [ctypes.c_int, D, D]
        lib.oracle_TDstressFS_C.restype = ctypes.c_int
        lib.oracle_TDstressEachSourceAtReceiver_C.argtypes = [ctypes.c_double] * 3 + [
            D, D, D, ctypes.c_size_t] + [ctypes.c_double] * 5 + [ctypes.c_int, D, D]
        lib.oracle_TDstressEachSourceAtReceiver_C.restype = ctypes.c_int
        lib.oracle_traction_block.argtypes = [D, D, D, L, ctypes.c_int64, L, ctypes.c_int64,
                                              ctypes.c_double, ctypes.c_double, D]
        lib.oracle_traction_block.restype = None
        lib.oracle_num_threads.restype = ctypes.c_int
        self.lib = lib

    def stress_points(self, pts, P1, P2, P3, Ss, Ds, Ts, mu, lam):
        pts = np.asarray(pts, dtype=np.float64)
        n = pts.shape[0]
        X, Y, Z = (np.ascontiguousarray(pts[:, k]) for k in range(3))
        P1, P2, P3 = (np.ascontiguousarray(v, dtype=np.float64) for v in (P1, P2, P3))
        s = np.zeros((n, 6))
        e = np.zeros((n, 6))
        rc = self.lib.oracle_TDstressFS_C(_p(X), _p(Y), _p(Z), n, _p(P1), _p(P2), _p(P3), Ss, Ds,
                                          Ts, mu, lam, 0, _p(s), _p(e))
        assert rc == 0
        return s, e

    def stress_sources(self, x, y, z, tri, Ss, Ds, Ts, mu, lam):
        tri = np.asarray(tri, dtype=np.float64)
        ns = tri.shape[0]
        P1, P2, P3 = (np.ascontiguousarray(tri[:, 3 * k:3 * k + 3]) for k in range(3))
        s = np.zeros((ns, 6))
        e = np.zeros((ns, 6))
        rc = self.lib.oracle_TDstressEachSourceAtReceiver_C(x, y, z, _p(P1), _p(P2), _p(P3), ns,
                                                            Ss, Ds, Ts, mu, lam, 0, _p(s), _p(e))
        assert rc == 0
        return s, e

    def traction_block(self, xg, eleVec, tri, rows, cols, mu, lam):
        """6 x m x n operators A1s, A2s, Bs, A1d, A2d, Bd (FP64)."""
        xg = np.ascontiguousarray(xg, dtype=np.float64)
        ev = np.ascontiguousarray(eleVec, dtype=np.float64)
        tri = np.ascontiguousarray(tri, dtype=np.float64)
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        cols = np.ascontiguousarray(cols, dtype=np.int64)
        out = np.zeros((6, len(rows), len(cols)))
        self.lib.oracle_traction_block(_p(xg), _p(ev), _p(tri), rows.ctypes.data_as(L), len(rows),
                                       cols.ctypes.data_as(L), len(cols), mu, lam, _p(out))
        return out

    def num_threads(self) -> int:
        return int(self.lib.oracle_num_threads())
