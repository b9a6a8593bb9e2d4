"""Triangular-dislocation (TDE) Green's functions on the GPU.

Two layers:

* ``TDstressFS_C`` / ``TDstressEachSourceAtReceiver_C`` -- numpy wrappers of
  the drop-in C symbols with the reference's argument meaning
  (src/TDstressFS_C.cpp:2048-2143; bound at src/Hmatrix.py:22-49).
* ``DeviceMesh`` -- device-resident triangles, per-triangle source records and
  receiver records (centroid + frame), with ``traction_block`` producing the six
  traction operators of any (rows x cols) block in one launch. This replaces
  the per-column / per-row ctypes loops of the reference evaluators
  (src/Hmatrix.py:318-511) and their Python traction projection (:295-313).
* ``TDdispFS`` / ``TDdispHS`` and ``DeviceMesh.displacement`` -- displacement
  of triangular dislocations (src/DH_greenfunction.py:99, :402) and its sum
  over the mesh with per-element slip (get_sumS, src/QDsim.py:25-53).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .mesh import FaultMesh, get_eleVec, receivers

OPS = ("A1s", "A2s", "Bs", "A1d", "A2d", "Bd")
_D = ctypes.POINTER(ctypes.c_double)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_D)


def TDstressFS_C(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, mu, lambda_, jud_halfspace=False):
    """Stress and strain (n x 6 each) at n receivers from one triangle, through
    the library's ``TDstressFS_C`` symbol. Raises on a library error (the
    symbol itself returns NaN outputs then)."""
    lib = _lib.load()
    X, Y, Z = (np.ascontiguousarray(v, dtype=np.float64).ravel() for v in (X, Y, Z))
    P1, P2, P3 = (np.ascontiguousarray(v, dtype=np.float64).ravel() for v in (P1, P2, P3))
    n = X.shape[0]
    stress = np.zeros(n * 6)
    strain = np.zeros(n * 6)
    lib.TDstressFS_C(_dp(X), _dp(Y), _dp(Z), n, _dp(P1), _dp(P2), _dp(P3), Ss, Ds, Ts, mu,
                     lambda_, bool(jud_halfspace), _dp(stress), _dp(strain))
    err = lib.pq3d_last_error()
    if err:
        raise RuntimeError(err.decode())
    return stress.reshape(n, 6), strain.reshape(n, 6)


def TDstressEachSourceAtReceiver_C(x, y, z, P1_flat, P2_flat, P3_flat, Ss, Ds, Ts, mu, lambda_,
                                   jud_halfspace=False):
    """Stress and strain (n_sources x 6) at one receiver from n triangles."""
    lib = _lib.load()
    P1, P2, P3 = (np.ascontiguousarray(v, dtype=np.float64).ravel() for v in
                  (P1_flat, P2_flat, P3_flat))
    ns = P1.shape[0] // 3
    stress = np.zeros(ns * 6)
    strain = np.zeros(ns * 6)
    lib.TDstressEachSourceAtReceiver_C(float(x), float(y), float(z), _dp(P1), _dp(P2), _dp(P3),
                                       ns, Ss, Ds, Ts, mu, lambda_, bool(jud_halfspace),
                                       _dp(stress), _dp(strain))
    err = lib.pq3d_last_error()
    if err:
        raise RuntimeError(err.decode())
    return stress.reshape(ns, 6), strain.reshape(ns, 6)


def _disp(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, nu, halfspace, device):
    dev = torch.device(device)
    X, Y, Z = (np.ascontiguousarray(v, dtype=np.float64).ravel() for v in (X, Y, Z))
    n = X.shape[0]
    if Y.shape[0] != n or Z.shape[0] != n:
        raise ValueError("X, Y and Z must have the same length")
    tri = np.concatenate([np.asarray(P, dtype=np.float64).ravel() for P in (P1, P2, P3)])
    if tri.shape[0] != 9:
        raise ValueError("P1, P2 and P3 must each hold 3 coordinates")
    rec = torch.from_numpy(source_records(tri)).to(dev)
    hrec = torch.from_numpy(hs_records(tri)).to(dev) if halfspace else None
    dX, dY, dZ = (torch.from_numpy(v).to(dev) for v in (X, Y, Z))
    out = torch.empty((n, 3), dtype=torch.float64, device=dev)
    _lib.call("pq3d_tde_disp_points", _lib.ptr(dX), _lib.ptr(dY), _lib.ptr(dZ), n,
              _lib.ptr(rec), _lib.ptr(hrec), float(Ss), float(Ds), float(Ts), float(nu),
              _lib.ptr(out), _lib.stream_handle())
    u = out.cpu().numpy()
    return u[:, 0].copy(), u[:, 1].copy(), u[:, 2].copy()


def TDdispFS(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, nu, device="cuda"):
    """Full-space displacement (ue, un, uv) at points (X, Y, Z) from one
    triangle with strike, dip and tensile slip (src/DH_greenfunction.py:99)."""
    return _disp(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, nu, False, device)


def TDdispHS(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, nu, device="cuda"):
    """Half-space displacement (ue, un, uv), free surface z = 0, points and
    triangle at z <= 0 (src/DH_greenfunction.py:402)."""
    return _disp(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, nu, True, device)


def source_records(tri: np.ndarray) -> np.ndarray:
    """(N, 32) per-triangle records, computed on the host by the library."""
    tri = np.ascontiguousarray(tri, dtype=np.float64).reshape(-1, 9)
    rec = np.zeros((tri.shape[0], 32))
    _lib.check(_lib.load().pq3d_tde_source_records(_dp(tri), tri.shape[0], _dp(rec)),
               "pq3d_tde_source_records")
    return rec


def hs_records(tri: np.ndarray) -> np.ndarray:
    """(N, 104) per-triangle half-space records (image triangle, side
    frames), computed on the host by the library."""
    tri = np.ascontiguousarray(tri, dtype=np.float64).reshape(-1, 9)
    h = np.zeros((tri.shape[0], 104))
    _lib.check(_lib.load().pq3d_tde_hs_records(_dp(tri), tri.shape[0], _dp(h)),
               "pq3d_tde_hs_records")
    return h


class DeviceMesh:
    """Device-resident geometry of one fault mesh.

    tri  (N, 9)  FP64 triangle vertices
    rec  (N, 32) FP64 per-source records (pq3d_tde_source_records)
    recv (N, 12) FP64 receiver records: centroid xg and frame ev1, ev2, ev3
    """

    def __init__(self, nodelst, elelst, mu, lambda_, device="cuda", jud_ele_order=False,
                 halfspace=False):
        self.mu = float(mu)
        self.lambda_ = float(lambda_)
        self.device = torch.device(device)
        self.nodelst = np.asarray(nodelst, dtype=np.float64)
        self.elelst = np.asarray(elelst, dtype=np.int64)
        e = self.elelst - 1
        tri = np.concatenate([self.nodelst[e[:, 0]], self.nodelst[e[:, 1]],
                              self.nodelst[e[:, 2]]], axis=1)
        self.eleVec, self.xg = get_eleVec(self.nodelst, self.elelst, jud_ele_order)
        self.n = tri.shape[0]
        self.tri = torch.from_numpy(np.ascontiguousarray(tri)).to(self.device)
        self.recv = torch.from_numpy(receivers(self.eleVec, self.xg)).to(self.device)
        self.rec = torch.from_numpy(source_records(tri)).to(self.device)
        # half space (free surface z = 0): second per-triangle record
        self.halfspace = bool(halfspace)
        self.hrec = (torch.from_numpy(hs_records(tri)).to(self.device) if self.halfspace
                     else None)

    @classmethod
    def from_mesh(cls, mesh: FaultMesh, mu, lambda_, device="cuda", halfspace=False,
                  jud_ele_order=False):
        return cls(mesh.nodelst, mesh.elelst, mu, lambda_, device, jud_ele_order=jud_ele_order,
                   halfspace=halfspace)

    def _index(self, idx):
        if idx is None:
            return None
        t = torch.as_tensor(np.asarray(idx, dtype=np.int64) if not torch.is_tensor(idx) else idx,
                            dtype=torch.int64)
        return t.to(self.device).contiguous()

    def traction_block(self, rows=None, cols=None, *, row0=0, m=None, col0=0, n=None, out=None,
                       dtype=torch.float64, order="ij", stream=None):
        """Six traction operators (6, m, n) of receivers ``rows`` x sources ``cols``.

        rows/cols: index arrays (numpy or device tensors), or None for the
        contiguous ranges [row0, row0+m) / [col0, col0+n).
        order "ij" stores out[op, i, j] row-major; "ji" stores out[op, j, i].
        """
        r = self._index(rows)
        c = self._index(cols)
        m = int(r.numel()) if r is not None else (self.n - row0 if m is None else int(m))
        n = int(c.numel()) if c is not None else (self.n - col0 if n is None else int(n))
        if r is None and not (0 <= row0 and row0 + m <= self.n):
            raise ValueError("row range outside the mesh")
        if c is None and not (0 <= col0 and col0 + n <= self.n):
            raise ValueError("column range outside the mesh")
        if r is not None and m and (int(r.min()) < 0 or int(r.max()) >= self.n):
            raise ValueError("row index outside the mesh")
        if c is not None and n and (int(c.min()) < 0 or int(c.max()) >= self.n):
            raise ValueError("column index outside the mesh")
        shape = (6, m, n) if order == "ij" else (6, n, m)
        if out is None:
            out = torch.empty(shape, dtype=dtype, device=self.device)
        if tuple(out.shape) != shape or out.dtype != dtype or not out.is_contiguous():
            raise ValueError("out has the wrong shape, dtype or layout")
        si, sj = (n, 1) if order == "ij" else (1, m)
        if dtype not in (torch.float64, torch.float32):
            raise ValueError("dtype must be float64 or float32")
        fn = "pq3d_tde_traction_block" + ("_hs" if self.halfspace else "") + \
            ("" if dtype == torch.float64 else "_f32")
        _lib.check_width("recv", self.recv)
        _lib.check_width("rec", self.rec)
        hs = (_lib.ptr(self.hrec),) if self.halfspace else ()
        _lib.call(fn, _lib.ptr(self.recv), _lib.ptr(self.rec), *hs, _lib.ptr(r), row0, m,
                  _lib.ptr(c), col0, n, self.mu, self.lambda_, _lib.ptr(out), si, sj, m * n,
                  out.numel(), _lib.stream_handle(stream))
        return out

    @property
    def nu(self) -> float:
        return 1.0 / (1.0 + self.lambda_ / self.mu) / 2.0

    def displacement(self, obs, ss, ds, ts=None, *, nu=None, out=None, stream=None):
        """Displacement (nobs, 3: east, north, up) at points ``obs`` (nobs x 3)
        summed over every element with slip ``ss``, ``ds`` and opening ``ts``
        (length-N arrays or scalars; ``ts`` None means no opening), in the
        mesh's full or half space (get_sumS, src/QDsim.py:25-53, sums unit
        dip slip with nu = 0.25). ``nu`` defaults to the mesh's Poisson ratio."""
        def dev(a, name):
            t = torch.as_tensor(a, dtype=torch.float64)
            if t.ndim == 0:
                t = t.expand(self.n)
            if t.shape != (self.n,):
                raise ValueError(f"{name} must be a scalar or have one value per element")
            return t.to(self.device).contiguous()
        p = torch.as_tensor(obs, dtype=torch.float64).to(self.device).contiguous()
        if p.ndim != 2 or p.shape[1] != 3:
            raise ValueError("obs must be (nobs, 3)")
        nobs = p.shape[0]
        s, d = dev(ss, "ss"), dev(ds, "ds")
        t = dev(ts, "ts") if ts is not None else None
        if out is None:
            out = torch.empty((nobs, 3), dtype=torch.float64, device=self.device)
        if tuple(out.shape) != (nobs, 3) or out.dtype != torch.float64 or not out.is_contiguous():
            raise ValueError("out must be a contiguous float64 (nobs, 3) tensor")
        _lib.check_width("rec", self.rec)
        _lib.call("pq3d_tde_disp_sum", _lib.ptr(p), nobs, _lib.ptr(self.rec),
                  _lib.ptr(self.hrec), self.n, _lib.ptr(s), _lib.ptr(d), _lib.ptr(t),
                  float(self.nu if nu is None else nu), _lib.ptr(out),
                  _lib.stream_handle(stream))
        return out
