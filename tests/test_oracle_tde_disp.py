"""Pins of the CPU displacement oracle (oracle/tde_disp_oracle.py): the
symmetric gradient of the displacement equals the strain of the stress
oracles (the full-space C oracle is pinned to the reference C++ golden
vectors, the half-space one to its golden vectors and the traction-free
surface); the displacement jumps by the Burgers vector across the triangle;
far-field decay is that of a point source (1/r^2)."""
import numpy as np
import pytest

from oracle import tde_hs_oracle
from oracle.tde_disp_oracle import _frame, td_disp_fs, td_disp_hs

MU = LAM = 32e9
NU = 0.25


def _tri(rng, depth):
    c = np.array([rng.uniform(-1e3, 1e3), rng.uniform(-1e3, 1e3), -depth])
    P = c + rng.uniform(-800, 800, (3, 3))
    P[:, 2] = np.minimum(P[:, 2], -50.0)
    return P


def _sym_grad(f, pts, h):
    G = np.zeros((len(pts), 3, 3))
    for j in range(3):
        d = np.zeros(3)
        d[j] = h
        G[:, :, j] = (f(pts + d) - f(pts - d)) / (2 * h)
    E = 0.5 * (G + G.transpose(0, 2, 1))
    return np.stack([E[:, 0, 0], E[:, 1, 1], E[:, 2, 2], E[:, 0, 1], E[:, 0, 2], E[:, 1, 2]], 1)


def _far_points(rng, P, n, zmax=None):
    c = P.mean(0)
    pts = c + rng.uniform(-3e3, 3e3, (4 * n, 3))
    d = np.linalg.norm(pts[:, None, :] - P[None], axis=2).min(1)
    keep = d > 600.0
    if zmax is not None:
        keep &= pts[:, 2] < zmax
    return pts[keep][:n]


@pytest.mark.parametrize("slip", [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.3, -0.7, 0.4)])
def test_full_space_gradient_is_the_oracle_strain(oracle, slip):
    rng = np.random.default_rng(1)
    for _ in range(3):
        P = _tri(rng, 5e3)
        pts = _far_points(rng, P, 40)
        f = lambda x: td_disp_fs(x, P[0], P[1], P[2], *slip, NU)
        E = _sym_grad(f, pts, 0.5)
        _, ref = oracle.stress_points(pts, P[0], P[1], P[2], *slip, MU, LAM)
        assert np.abs(E - ref).max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("vertical", [False, True])
def test_half_space_gradient_is_the_oracle_strain(oracle, vertical):
    rng = np.random.default_rng(2)
    for _ in range(3):
        P = _tri(rng, 2.5e3)
        if vertical:                 # a vertical side: P1 above P0
            P[1, :2] = P[0, :2]
        pts = _far_points(rng, P, 40, zmax=-100.0)
        f = lambda x: td_disp_hs(x, P[0], P[1], P[2], 0.6, -0.8, 0.2, NU)
        E = _sym_grad(f, pts, 0.5)
        _, ref = tde_hs_oracle.stress_hs(oracle, pts, P, 0.6, -0.8, 0.2, MU, LAM, stale=False)
        assert np.abs(E - ref).max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("hs", [False, True])
def test_jump_across_the_triangle_is_the_burgers_vector(hs):
    rng = np.random.default_rng(3)
    P = _tri(rng, 4e3)
    vn, vs, vd = _frame(*P)
    Ss, Ds, Ts = 0.7, -0.4, 0.25
    b = vs * Ss + vd * Ds + vn * Ts
    w = rng.dirichlet([2, 2, 2], 10)               # interior barycentric points
    x0 = w @ P
    f = td_disp_hs if hs else td_disp_fs
    up = f(x0 + 1e-5 * vn, *P, Ss, Ds, Ts, NU)
    um = f(x0 - 1e-5 * vn, *P, Ss, Ds, Ts, NU)
    # crossing the dislocation along its normal adds the Burgers vector
    assert np.abs((up - um) - b).max() < 1e-6


def test_far_field_decays_as_a_point_source():
    rng = np.random.default_rng(4)
    P = _tri(rng, 5e3)
    c = P.mean(0)
    d = rng.normal(size=(20, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    u1 = td_disp_fs(c + 2e5 * d, *P, 1.0, 0.0, 0.0, NU)
    u2 = td_disp_fs(c + 4e5 * d, *P, 1.0, 0.0, 0.0, NU)
    ratio = np.linalg.norm(u2, axis=1) / np.linalg.norm(u1, axis=1)
    assert np.allclose(ratio, 0.25, rtol=2e-2)
