"""The device TDE arithmetic, compiled for the host, reproduces the
reference C++ bit for bit on every golden vector (including the near-singular
points where the reference is non-finite), and the fused strike+dip traction
path reproduces two separate reference evaluations plus the projection."""
import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest

from pyquake3d_amd.mesh import get_eleVec, receivers

REPO = Path(__file__).resolve().parents[1]
D = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def host_lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("hostbuild") / "libtde_host.so"
    src = REPO / "tests" / "host_build" / "tde_host.cpp"
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared",
           f"-I{REPO / 'pyquake3d_amd' / 'csrc'}", str(src), "-o", str(out)]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    lib = ctypes.CDLL(str(out))
    lib.host_tde_points.argtypes = [D, ctypes.c_int64, D] + [ctypes.c_double] * 5 + [D, D]
    lib.host_tde_block.argtypes = [D, ctypes.c_int64, D, ctypes.c_int64, ctypes.c_double,
                                   ctypes.c_double, D]
    lib.host_tde_points_hs.argtypes = [D, ctypes.c_int64, D] + [ctypes.c_double] * 5 + [D, D]
    lib.host_tde_block_hs.argtypes = [D, ctypes.c_int64, D, ctypes.c_int64, ctypes.c_double,
                                      ctypes.c_double, D]
    lib.host_tde_disp_points.argtypes = [D, ctypes.c_int64, D] + [ctypes.c_double] * 4 + \
        [ctypes.c_int, D]
    return lib


def p(a):
    return a.ctypes.data_as(D)


def test_host_build_bitwise_reference_points(host_lib, golden):
    g = golden["g1_points"]
    n = len(g["pts"])
    S = np.zeros((n, 6))
    E = np.zeros((n, 6))
    for i in range(n):
        s = np.zeros(6)
        e = np.zeros(6)
        host_lib.host_tde_points(p(np.ascontiguousarray(g["pts"][i])), 1,
                                 p(np.ascontiguousarray(g["tri"][i])), *g["bv"][i], *g["ml"][i],
                                 p(s), p(e))
        S[i], E[i] = s, e
    R = g["stress"]
    fin = np.isfinite(R).all(1)
    assert (np.isfinite(S).all(1) == fin).all()
    assert np.array_equal(S[fin], R[fin])
    assert np.array_equal(E[fin], g["strain"][fin])


@pytest.mark.parametrize("tag", ["planar", "undul"])
def test_host_build_bitwise_reference_blocks(host_lib, golden, tag):
    g = golden["g3_blocks"]
    nodes, ele, K = g[f"{tag}_nodes"], g[f"{tag}_elelst"], g[f"{tag}_K"]
    ev, xg = get_eleVec(nodes, ele)
    e = ele - 1
    tri = np.ascontiguousarray(np.concatenate([nodes[e[:, 0]], nodes[e[:, 1]], nodes[e[:, 2]]], 1))
    recv = receivers(ev, xg)
    n = len(ele)
    out = np.zeros((6, n, n))
    host_lib.host_tde_block(p(recv), n, p(tri), n, 32e9, 32e9, p(out))
    assert np.array_equal(out, K)


def test_host_build_halfspace_points(host_lib, golden):
    """Half space: the forward-mode correction strain agrees with the
    reference's expanded formula to 1e-10 per point (not bitwise: another
    evaluation order), on buried, vertical (vertical side first, second and
    third), surface-breaking and dipping triangles, receivers on and below
    the free surface."""
    g = golden["g4_halfspace"]
    n = len(g["pts"])
    S = np.zeros((n, 6))
    E = np.zeros((n, 6))
    for i in range(n):
        s = np.zeros(6)
        e = np.zeros(6)
        host_lib.host_tde_points_hs(p(np.ascontiguousarray(g["pts"][i])), 1,
                                    p(np.ascontiguousarray(g["tri"][i])), *g["bv"][i],
                                    *g["ml"][i], p(s), p(e))
        S[i], E[i] = s, e
    for got, ref in ((S, g["stress"]), (E, g["strain"])):
        den = np.maximum(np.abs(ref).max(1), 1e-300)
        assert (np.abs(got - ref).max(1) / den).max() < 1e-10


def test_host_build_halfspace_block(host_lib, golden):
    g = golden["g4_halfspace"]
    nodes, ele, K = g["blk_nodes"], g["blk_elelst"], g["blk_K"]
    ev, xg = get_eleVec(nodes, ele)
    e = ele - 1
    tri = np.ascontiguousarray(np.concatenate([nodes[e[:, 0]], nodes[e[:, 1]], nodes[e[:, 2]]],
                                              axis=1))
    recv = np.ascontiguousarray(receivers(ev, xg))
    n = len(tri)
    out = np.zeros((6, n, n))
    host_lib.host_tde_block_hs(p(recv), n, p(tri), n, 32e9, 32e9, p(out))
    scale = np.abs(K).max(axis=(0, 2))[None, :, None]
    assert (np.abs(out - K) / scale).max() < 1e-10


@pytest.mark.parametrize("hs", [0, 1])
def test_host_build_displacement_matches_oracle(host_lib, hs):
    """The device displacement arithmetic (full and half space) against the
    displacement oracle, to 1e-10 of each point's largest component plus
    1e-13 |b| (far from a small triangle the displacement is a small
    difference of terms of size |b|, so rounding there is relative to |b|),
    on random buried triangles (every fifth with a vertical side) and points
    down to the free surface."""
    from oracle.tde_disp_oracle import td_disp_fs, td_disp_hs
    rng = np.random.default_rng(11 + hs)
    for t in range(20):
        c = np.array([rng.uniform(-1e3, 1e3), rng.uniform(-1e3, 1e3), -rng.uniform(200, 3000)])
        P = c + rng.uniform(-800, 800, (3, 3))
        P[:, 2] = np.minimum(P[:, 2], -10.0)
        if t % 5 == 0:
            P[1, :2] = P[0, :2]
        pts = c + rng.uniform(-3e3, 3e3, (100, 3))
        pts[:, 2] = np.minimum(pts[:, 2], 0.0)
        b = rng.normal(size=3)
        out = np.zeros((100, 3))
        host_lib.host_tde_disp_points(p(np.ascontiguousarray(pts)), 100,
                                      p(np.ascontiguousarray(P.ravel())), *b, 0.25, hs, p(out))
        ref = (td_disp_hs if hs else td_disp_fs)(pts, P[0], P[1], P[2], *b, 0.25)
        tol = 1e-10 * np.abs(ref).max(1) + 1e-13 * np.linalg.norm(b)
        assert (np.abs(out - ref).max(1) <= tol).all()
