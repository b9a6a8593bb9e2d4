"""Write the golden vectors under tests/golden/ from the reference's own C++.

Runs only where the reference source exists (it is compiled from source by
oracle/Makefile into oracle/_ref/libTDstressFS_ref.so, never the shipped
binary). The fixtures it writes are data: inputs and the reference outputs.

  g1_points.npz   TDstressFS_C (src/TDstressFS_C.cpp:2048): n receivers x one
                  triangle, for a set of triangles and point families
  g2_sources.npz  TDstressEachSourceAtReceiver_C (src/TDstressFS_C.cpp:2098):
                  one receiver x n triangles
  g3_blocks.npz   the six traction operators of two small meshes: the
                  reference C++ per source column (as src/Hmatrix.py:425-511
                  calls it) composed with the projection of
                  src/Hmatrix.py:295-313
  g4_halfspace.npz TDstressFS_C with the half-space flag
                  (src/TDstressFS_C.cpp:1911-1999): buried, vertical (a
                  vertical side in each of the three positions),
                  surface-breaking and dipping triangles; receivers below and
                  on the free surface; and the six traction operators of a
                  small surface-breaking mesh

Usage: python tests/golden/make_golden.py [g1 g2 g3 g4]   (default: all)
"""
from __future__ import annotations

import ctypes
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]
REF_LIB = REPO / "oracle" / "_ref" / "libTDstressFS_ref.so"

sys.path.insert(0, str(REPO))
from pyquake3d_amd.mesh import get_eleVec, planar_fault, undulating_fault  # noqa: E402

D = ctypes.POINTER(ctypes.c_double)


def load_ref():
    if not REF_LIB.exists():
        subprocess.run(["make", "-C", str(REPO / "oracle"), "ref"], check=True)
    lib = ctypes.CDLL(str(REF_LIB))
    lib.TDstressFS_C.argtypes = [D, D, D, ctypes.c_size_t, D, D, D] + [ctypes.c_double] * 5 + [
        ctypes.c_bool, D, D]
    lib.TDstressEachSourceAtReceiver_C.argtypes = [ctypes.c_double] * 3 + [D, D, D,
                                                                          ctypes.c_size_t] + [
        ctypes.c_double] * 5 + [ctypes.c_bool, D, D]
    return lib


def p(a):
    return a.ctypes.data_as(D)


def ref_points(lib, X, Y, Z, P1, P2, P3, Ss, Ds, Ts, mu, lam, halfspace=False):
    n = len(X)
    s = np.zeros(6 * n)
    e = np.zeros(6 * n)
    X, Y, Z = (np.ascontiguousarray(v, dtype=np.float64) for v in (X, Y, Z))
    P1, P2, P3 = (np.ascontiguousarray(v, dtype=np.float64) for v in (P1, P2, P3))
    lib.TDstressFS_C(p(X), p(Y), p(Z), n, p(P1), p(P2), p(P3), Ss, Ds, Ts, mu, lam, halfspace,
                     p(s), p(e))
    return s.reshape(n, 6), e.reshape(n, 6)


def ref_sources(lib, x, y, z, P1f, P2f, P3f, Ss, Ds, Ts, mu, lam):
    ns = P1f.shape[0]
    s = np.zeros(6 * ns)
    e = np.zeros(6 * ns)
    P1f, P2f, P3f = (np.ascontiguousarray(v, dtype=np.float64).ravel() for v in (P1f, P2f, P3f))
    lib.TDstressEachSourceAtReceiver_C(x, y, z, p(P1f), p(P2f), p(P3f), ns, Ss, Ds, Ts, mu, lam,
                                       False, p(s), p(e))
    return s.reshape(ns, 6), e.reshape(ns, 6)


def point_families(rng, P):
    """Receiver points around triangle P (3x3) covering the edge cases."""
    c = P.mean(0)
    L = np.linalg.norm(P[1] - P[0])
    nrm = np.cross(P[1] - P[0], P[2] - P[0])
    nrm /= np.linalg.norm(nrm)
    pts = []
    # far / mid / near, random directions, distances 1e-3..1e2 edge lengths
    d = rng.normal(size=(60, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = L * 10 ** rng.uniform(-3, 2, size=(60, 1))
    pts.append(c + d * r)
    # near vertices and edges, 1e-8..1e-1 edge lengths off
    for k in range(3):
        v = P[k]
        w = P[(k + 1) % 3]
        for s in 10.0 ** np.arange(-8, 0):
            off = rng.normal(size=3)
            off /= np.linalg.norm(off)
            pts.append((v + s * L * off)[None])
            pts.append((0.5 * (v + w) + s * L * off)[None])
    # in-plane points (inside, outside, on side extensions)
    u = P[1] - P[0]
    w = P[2] - P[0]
    ab = rng.uniform(-1.0, 2.0, size=(30, 2))
    pts.append(P[0] + ab[:, :1] * u + ab[:, 1:] * w)
    # slightly off-plane
    pts.append(P[0] + ab[:10, :1] * u + ab[:10, 1:] * w + 1e-6 * L * nrm)
    # self-term: the centroid
    pts.append(c[None])
    return np.concatenate(pts, axis=0)


def make_g1(lib, rng):
    tris = []
    # random triangles of varied shape and orientation
    for _ in range(12):
        P = rng.normal(size=(3, 3)) * 1000.0
        tris.append(P)
    # vertical triangles in the plane y=0 with exact coordinates (coplanar points give x=0 exactly)
    tris.append(np.array([[0.0, 0.0, -1000.0], [0.0, 0.0, 0.0], [1000.0, 0.0, -1000.0]]))
    tris.append(np.array([[-500.0, 0.0, -2000.0], [500.0, 0.0, -2000.0], [0.0, 0.0, -1000.0]]))
    # dipping triangle
    tris.append(np.array([[0.0, 0.0, -3000.0], [2000.0, 0.0, -3000.0], [1000.0, 1500.0, -1500.0]]))
    recs = []
    for ti, P in enumerate(tris):
        pts = point_families(rng, P)
        if ti in (12, 13):
            # exact in-plane points: side midpoints (mode 0 -> zero) and side extensions
            ex = [0.5 * (P[0] + P[1]), 0.5 * (P[1] + P[2]), 0.5 * (P[0] + P[2]),
                  2 * P[1] - P[0], 2 * P[2] - P[1], P[0] + 3 * (P[2] - P[0]),
                  np.array([300.0, 0.0, -500.0]), np.array([-700.0, 0.0, 250.0])]
            pts = np.concatenate([pts, np.array(ex)], axis=0)
        for (Ss, Ds, Ts) in ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (0.7, -0.4, 0.25)):
            for (mu, lam) in ((32e9, 32e9), (30e9, 20e9), (30e9, 45e9)):
                if (mu, lam) != (32e9, 32e9) and (Ss, Ds, Ts) != (0.7, -0.4, 0.25):
                    continue
                s, e = ref_points(lib, pts[:, 0], pts[:, 1], pts[:, 2], P[0], P[1], P[2], Ss, Ds,
                                  Ts, mu, lam)
                recs.append(dict(tri=np.repeat(P.reshape(1, 9), len(pts), 0), pts=pts,
                                 bv=np.repeat([[Ss, Ds, Ts]], len(pts), 0),
                                 ml=np.repeat([[mu, lam]], len(pts), 0), stress=s, strain=e,
                                 group=np.full(len(pts), ti)))
    out = {k: np.concatenate([r[k] for r in recs], axis=0) for k in recs[0]}
    # horizontal triangle: the reference divides by a zero norm (non-finite output)
    Ph = np.array([[0.0, 0.0, -1000.0], [1000.0, 0.0, -1000.0], [0.0, 1000.0, -1000.0]])
    hp = np.array([[300.0, 300.0, -500.0], [2000.0, -100.0, -1500.0]])
    hs, he = ref_points(lib, hp[:, 0], hp[:, 1], hp[:, 2], Ph[0], Ph[1], Ph[2], 1.0, 0.0, 0.0,
                        32e9, 32e9)
    out.update(h_tri=Ph.reshape(1, 9), h_pts=hp, h_stress=hs, h_strain=he)
    return out


def make_g2(lib, rng):
    out = {}
    mesh = undulating_fault(length=6e3, depth=3e3, nx=6, nz=3, amplitude=300, wavelength=5e3,
                            seed=3)
    tri = mesh.triangles()
    eleVec, xg = get_eleVec(mesh.nodelst, mesh.elelst)
    recv = np.concatenate([xg[:5], np.array([[100.0, 2000.0, -1500.0], [0.0, 0.0, -10.0]])], 0)
    S, E = [], []
    for (x, y, z) in recv:
        for (Ss, Ds, Ts) in ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.3, 0.5, -0.2)):
            s, e = ref_sources(lib, x, y, z, tri[:, 0:3], tri[:, 3:6], tri[:, 6:9], Ss, Ds, Ts,
                               32e9, 32e9)
            S.append(s)
            E.append(e)
    out.update(tri=tri, recv=recv, bvs=np.array([[1.0, 0, 0], [0, 1.0, 0], [0.3, 0.5, -0.2]]),
               stress=np.stack(S).reshape(len(recv), 3, -1, 6),
               strain=np.stack(E).reshape(len(recv), 3, -1, 6))
    return out


def traction(stress, f):
    """src/Hmatrix.py:295-313 for one receiver frame f (9,) and stresses (k, 6)."""
    s = stress
    t1 = s[:, 0] * f[6] + s[:, 3] * f[7] + s[:, 4] * f[8]
    t2 = s[:, 3] * f[6] + s[:, 1] * f[7] + s[:, 5] * f[8]
    tn = s[:, 4] * f[6] + s[:, 5] * f[7] + s[:, 2] * f[8]
    return np.stack([t1 * f[0] + t2 * f[1] + tn * f[2], t1 * f[3] + t2 * f[4] + tn * f[5],
                     t1 * f[6] + t2 * f[7] + tn * f[8]], axis=1)


def ref_block(lib, mesh, mu, lam, halfspace=False):
    tri = mesh.triangles()
    eleVec, xg = get_eleVec(mesh.nodelst, mesh.elelst)
    n = mesh.n
    K = np.zeros((6, n, n))
    for j in range(n):
        for op0, (Ss, Ds) in ((0, (1.0, 0.0)), (3, (0.0, 1.0))):
            s, _ = ref_points(lib, xg[:, 0], xg[:, 1], xg[:, 2], tri[j, 0:3], tri[j, 3:6],
                              tri[j, 6:9], Ss, Ds, 0.0, mu, lam, halfspace)
            for i in range(n):
                t = traction(s[i:i + 1], eleVec[i])[0]
                K[op0:op0 + 3, i, j] = t
    return K, eleVec, xg, tri


def make_g3(lib):
    out = {}
    for tag, mesh in (("planar", planar_fault(length=10e3, depth=5e3, nx=10, nz=5, seed=0)),
                      ("undul", undulating_fault(length=12e3, depth=6e3, nx=12, nz=6,
                                                 amplitude=300, wavelength=5e3, seed=1))):
        K, eleVec, xg, tri = ref_block(lib, mesh, 32e9, 32e9)
        out[f"{tag}_nodes"] = mesh.nodelst
        out[f"{tag}_elelst"] = mesh.elelst
        out[f"{tag}_K"] = K
    return out


def make_g4(lib):
    rng = np.random.default_rng(4404)
    tris = []
    for _ in range(4):   # buried, random orientation
        P = rng.normal(size=(3, 3)) * 800.0
        P[:, 2] = -np.abs(P[:, 2]) - 1500.0
        tris.append(P)
    # vertical triangles of a ruled mesh: the vertical side first, second, third
    tris.append(np.array([[0.0, 0.0, -1000.0], [0.0, 0.0, -2000.0], [1000.0, 0.0, -2000.0]]))
    tris.append(np.array([[1000.0, 0.0, -2000.0], [0.0, 0.0, -1000.0], [0.0, 0.0, -2000.0]]))
    tris.append(np.array([[0.0, 0.0, -2000.0], [1000.0, 0.0, -2000.0], [0.0, 0.0, -1000.0]]))
    # vertical surface-breaking: a side along z = 0
    tris.append(np.array([[0.0, 0.0, 0.0], [1000.0, 0.0, -1000.0], [1000.0, 0.0, 0.0]]))
    # dipping, one vertex at the surface
    tris.append(np.array([[0.0, 0.0, 0.0], [1500.0, 300.0, -1200.0], [-200.0, 1200.0, -800.0]]))
    # undulating-strand triangle (vertical, not in a coordinate plane)
    tris.append(np.array([[200.0, 150.0, -3000.0], [200.0, 150.0, -2000.0],
                          [1100.0, -120.0, -3000.0]]))
    recs = []
    for ti, P in enumerate(tris):
        c = P.mean(0)
        L = np.linalg.norm(P[1] - P[0])
        d = rng.normal(size=(40, 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        pts = c + d * (L * 10 ** rng.uniform(-1.5, 1.0, size=(40, 1)))
        pts[:, 2] = -np.abs(pts[:, 2])                      # below the surface
        surf = np.stack([rng.uniform(-3e3, 3e3, 12), rng.uniform(-3e3, 3e3, 12),
                         np.zeros(12)], 1)                 # on the free surface
        pts = np.concatenate([pts, surf], 0)
        for (Ss, Ds, Ts) in ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (0.6, -0.3, 0.2)):
            for (mu, lam) in ((32e9, 32e9), (30e9, 45e9)):
                if (mu, lam) != (32e9, 32e9) and (Ss, Ds, Ts) != (0.6, -0.3, 0.2):
                    continue
                st, sn = ref_points(lib, pts[:, 0], pts[:, 1], pts[:, 2], P[0], P[1], P[2], Ss,
                                    Ds, Ts, mu, lam, halfspace=True)
                recs.append(dict(tri=np.repeat(P.reshape(1, 9), len(pts), 0), pts=pts,
                                 bv=np.repeat([[Ss, Ds, Ts]], len(pts), 0),
                                 ml=np.repeat([[mu, lam]], len(pts), 0), stress=st, strain=sn,
                                 group=np.full(len(pts), ti)))
    out = {k: np.concatenate([r[k] for r in recs], axis=0) for k in recs[0]}
    mesh = planar_fault(length=6e3, depth=3e3, nx=6, nz=3, seed=2)
    K, eleVec, xg, tri = ref_block(lib, mesh, 32e9, 32e9, halfspace=True)
    out.update(blk_nodes=mesh.nodelst, blk_elelst=mesh.elelst, blk_K=K)
    return out


def main():
    lib = load_ref()
    which = sys.argv[1:] or ["g1", "g2", "g3", "g4"]
    rng = np.random.default_rng(20240601)
    g1 = make_g1(lib, rng)        # g1 and g2 draw from one generator, in this order
    g2 = make_g2(lib, rng)
    if "g1" in which:
        np.savez_compressed(HERE / "g1_points.npz", **g1)
    if "g2" in which:
        np.savez_compressed(HERE / "g2_sources.npz", **g2)
    if "g3" in which:
        np.savez_compressed(HERE / "g3_blocks.npz", **make_g3(lib))
    if "g4" in which:
        np.savez_compressed(HERE / "g4_halfspace.npz", **make_g4(lib))
    for f in sorted(HERE.glob("g*.npz")):
        print(f.name, f.stat().st_size)


if __name__ == "__main__":
    main()
