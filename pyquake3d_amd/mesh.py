"""Synthetic fault meshes and per-element frames.

The reference reads gmsh meshes (src/readmsh.py:12-68); the workloads here are
synthetic, so this module generates the same kind of input: node coordinates in
metres and 1-based triangle connectivity (``elelst``), as ``read_mshV2`` returns
them.

``get_eleVec`` restates src/readmsh.py:75-134 (element frames ev1, ev2, ev3 and
centroids) in vectorised numpy.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class FaultMesh:
    nodelst: np.ndarray  # (M, 3) float64, metres
    elelst: np.ndarray   # (N, 3) int64, 1-based node ids
    name: str = ""

    @property
    def n(self) -> int:
        return int(self.elelst.shape[0])

    def triangles(self) -> np.ndarray:
        """(N, 9) float64: P1, P2, P3 of each element."""
        e = self.elelst - 1
        return np.ascontiguousarray(
            np.concatenate([self.nodelst[e[:, 0]], self.nodelst[e[:, 1]], self.nodelst[e[:, 2]]],
                           axis=1))


def _ruled_fault(xs: np.ndarray, zs: np.ndarray, yfun, seed: int, name: str) -> FaultMesh:
    """Vertical ruled surface y = yfun(x) over a tensor grid (xs, zs).

    Each rectangle is split on a diagonal whose direction is drawn from
    ``seed``. Node order gives every triangle a normal with positive y
    component for a planar fault.
    """
    nx, nz = len(xs) - 1, len(zs) - 1
    X, Z = np.meshgrid(xs, zs, indexing="ij")  # (nx+1, nz+1)
    Y = yfun(X)
    nodes = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1).astype(np.float64)
    nid = np.arange((nx + 1) * (nz + 1)).reshape(nx + 1, nz + 1)
    a = nid[:-1, :-1].ravel()   # (x0, z0)
    b = nid[1:, :-1].ravel()    # (x1, z0)
    c = nid[1:, 1:].ravel()     # (x1, z1)
    d = nid[:-1, 1:].ravel()    # (x0, z1)
    rng = np.random.default_rng(seed)
    diag = rng.integers(0, 2, size=a.shape[0]).astype(bool)
    # diagonal a-c: (a, c, b), (a, d, c); diagonal b-d: (a, d, b), (b, d, c)
    t1 = np.where(diag[:, None], np.stack([a, c, b], 1), np.stack([a, d, b], 1))
    t2 = np.where(diag[:, None], np.stack([a, d, c], 1), np.stack([b, d, c], 1))
    tris = np.empty((2 * a.shape[0], 3), dtype=np.int64)
    tris[0::2] = t1
    tris[1::2] = t2
    return FaultMesh(nodes, tris + 1, name)


def planar_fault(length=40e3, depth=20e3, nx=45, nz=22, seed=0) -> FaultMesh:
    """Vertical strike-slip fault in the plane y=0, x in [-L/2, L/2], z in [-depth, 0]."""
    xs = np.linspace(-length / 2, length / 2, nx + 1)
    zs = np.linspace(-depth, 0.0, nz + 1)
    return _ruled_fault(xs, zs, lambda x: np.zeros_like(x), seed, f"planar_{2 * nx * nz}")


def undulating_fault(length=100e3, depth=25e3, nx=200, nz=50, amplitude=300.0,
                     wavelength=5e3, seed=0) -> FaultMesh:
    """Vertical ruled fault y(x) = A sin(2 pi x / lambda + phi), phi ~ U[0, 2 pi) from seed."""
    phi = np.random.default_rng(seed + 7919).uniform(0.0, 2 * np.pi)
    xs = np.linspace(-length / 2, length / 2, nx + 1)
    zs = np.linspace(-depth, 0.0, nz + 1)
    return _ruled_fault(xs, zs, lambda x: amplitude * np.sin(2 * np.pi * x / wavelength + phi),
                        seed, f"undulating_{2 * nx * nz}")


def get_eleVec(nodelst: np.ndarray, elelst: np.ndarray, jud_ele_order: bool = False):
    """Element frames and centroids (src/readmsh.py:75-134).

    Returns ``eleVec`` (N, 9) = ev1, ev2, ev3 per row and ``xg`` (N, 3).
    ev3 is the unit normal (b-a) x (c-a) (or (c-a) x (b-a) when
    ``jud_ele_order``), ev1 = (ev32, -ev31, 0)/|.| unless |ev33| == 1 where it
    is (1, 0, 0), and ev2 = ev3 x ev1.
    """
    e = np.asarray(elelst, dtype=np.int64) - 1
    xa, xb, xc = nodelst[e[:, 0]], nodelst[e[:, 1]], nodelst[e[:, 2]]
    xg = np.stack([(xa[:, k] + xb[:, k] + xc[:, k]) / 3.0 for k in range(3)], axis=1)
    vba, vca = xb - xa, xc - xa
    if jud_ele_order:
        ev3 = np.cross(vca, vba)
    else:
        ev3 = np.cross(vba, vca)
    ev3 = ev3 / np.sqrt((ev3 ** 2).sum(1))[:, None]
    ev1 = np.zeros_like(ev3)
    tilted = np.abs(ev3[:, 2]) < 1
    ev1[:, 0] = np.where(tilted, ev3[:, 1], 1.0)
    ev1[:, 1] = np.where(tilted, -ev3[:, 0], 0.0)
    rr = np.sqrt(ev1[:, 0] ** 2 + ev1[:, 1] ** 2)
    ev1[:, 0] /= rr
    ev1[:, 1] /= rr
    ev2 = np.stack([ev3[:, 1] * ev1[:, 2] - ev3[:, 2] * ev1[:, 1],
                    ev3[:, 2] * ev1[:, 0] - ev3[:, 0] * ev1[:, 2],
                    ev3[:, 0] * ev1[:, 1] - ev3[:, 1] * ev1[:, 0]], axis=1)
    eleVec = np.ascontiguousarray(np.concatenate([ev1, ev2, ev3], axis=1))
    return eleVec, np.ascontiguousarray(xg)


def receivers(eleVec: np.ndarray, xg: np.ndarray) -> np.ndarray:
    """(N, 12) float64 receiver records: centroid (3) then ev1, ev2, ev3."""
    return np.ascontiguousarray(np.concatenate([xg, eleVec], axis=1))
