"""GPU parity of the displacement kernels (TDdispFS / TDdispHS and the summed
displacement of a mesh) against the displacement oracle. Tolerance per point:
max_c |du_c| <= 1e-10 max_c |u_c| + 1e-13 |b| (|b|: the slip magnitude, summed
over the sources for a mesh sum); far from a small triangle the displacement
is a small difference of terms of size |b|, so rounding there scales with |b|.
"""
import numpy as np
import pytest
import torch

from oracle.tde_disp_oracle import td_disp_fs, td_disp_hs
from pyquake3d_amd import _lib, tde
from pyquake3d_amd.mesh import planar_fault

pytestmark = pytest.mark.gpu
MU = LAM = 32e9
NU = 0.25


def _tri(rng, t):
    c = np.array([rng.uniform(-1e3, 1e3), rng.uniform(-1e3, 1e3), -rng.uniform(200, 3000)])
    P = c + rng.uniform(-800, 800, (3, 3))
    P[:, 2] = np.minimum(P[:, 2], -10.0)
    if t % 4 == 0:
        P[1, :2] = P[0, :2]          # a vertical side
    return c, P


def _close(got, ref, bnorm):
    tol = 1e-10 * np.abs(ref).max(1) + 1e-13 * bnorm
    return np.abs(got - ref).max(1) <= tol


@pytest.mark.parametrize("hs", [False, True])
def test_points_match_oracle(gpu, pq3d_lib, hs):
    rng = np.random.default_rng(5 + hs)
    fn, ref_fn = (tde.TDdispHS, td_disp_hs) if hs else (tde.TDdispFS, td_disp_fs)
    for t in range(12):
        c, P = _tri(rng, t)
        n = 300 + 37 * t                  # not a multiple of the workgroup
        pts = c + rng.uniform(-3e3, 3e3, (n, 3))
        pts[:, 2] = np.minimum(pts[:, 2], 0.0)
        pts[:8, 2] = 0.0                  # on the free surface
        b = rng.normal(size=3)
        ue, un, uv = fn(pts[:, 0], pts[:, 1], pts[:, 2], P[0], P[1], P[2], *b, NU)
        got = np.stack([ue, un, uv], 1)
        ref = ref_fn(pts, P[0], P[1], P[2], *b, NU)
        assert _close(got, ref, np.linalg.norm(b)).all(), (t, hs)


def test_points_in_plane_sides_nan_like_oracle(gpu, pq3d_lib):
    """Points on a side of the triangle in its own plane are NaN (mode 0);
    points in the plane inside and outside the triangle are finite."""
    P = np.array([[0.0, 0.0, -1000.0], [1000.0, 0.0, -1500.0], [0.0, 800.0, -1200.0]])
    mid = 0.5 * (P[0] + P[1])
    inside = P.mean(0)
    outside = P[0] + 1.5 * (P[1] - P[0])       # on the side's line, past the vertex
    pts = np.array([mid, inside, outside, inside + [0.0, 0.0, 10.0]])
    ue, un, uv = tde.TDdispFS(pts[:, 0], pts[:, 1], pts[:, 2], P[0], P[1], P[2], 1.0, 0.5, 0.0,
                              NU)
    got = np.stack([ue, un, uv], 1)
    ref = td_disp_fs(pts, P[0], P[1], P[2], 1.0, 0.5, 0.0, NU)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    fin = np.isfinite(ref).all(1)
    assert not fin[0] and fin[1:].all()
    assert _close(got[fin], ref[fin], np.sqrt(1.25)).all()


def _dipping_mesh(dip_deg=60.0):
    m = planar_fault(length=4e3, depth=2e3, nx=6, nz=4)
    nodes = m.nodelst.copy()
    nodes[:, 1] = -nodes[:, 2] / np.tan(np.radians(dip_deg))
    return nodes, m.elelst


@pytest.mark.parametrize("hs", [False, True])
@pytest.mark.parametrize("geom", ["vertical", "dipping"])
def test_mesh_sum_matches_oracle(gpu, pq3d_lib, hs, geom):
    if geom == "vertical":
        m = planar_fault(length=4e3, depth=2e3, nx=6, nz=4)
        nodes, ele = m.nodelst, m.elelst
    else:
        nodes, ele = _dipping_mesh()
    dm = tde.DeviceMesh(nodes, ele, MU, LAM, device=gpu, halfspace=hs)
    rng = np.random.default_rng(17)
    nobs = 100
    obs = np.stack([rng.uniform(-4e3, 4e3, nobs), rng.uniform(-3e3, 3e3, nobs),
                    -rng.uniform(0, 3e3, nobs)], 1)
    obs[:40, 2] = 0.0
    ss, ds, ts = rng.normal(size=(3, dm.n))
    got = dm.displacement(obs, ss, ds, ts).cpu().numpy()
    ref_fn = td_disp_hs if hs else td_disp_fs
    e = ele - 1
    ref = np.zeros((nobs, 3))
    for j in range(dm.n):
        ref += ref_fn(obs, nodes[e[j, 0]], nodes[e[j, 1]], nodes[e[j, 2]], ss[j], ds[j], ts[j],
                      NU)
    bsum = np.sqrt(ss ** 2 + ds ** 2 + ts ** 2).sum()
    assert _close(got, ref, bsum).all()
    again = dm.displacement(obs, ss, ds, ts).cpu().numpy()
    assert np.array_equal(got, again)               # deterministic


def test_mesh_sum_unit_dip_equals_per_triangle_sum(gpu, pq3d_lib):
    """The reference's get_sumS case: unit dip slip on every element, nu 0.25,
    half space; equal to the sum of TDdispHS over the elements."""
    nodes, ele = _dipping_mesh(45.0)
    dm = tde.DeviceMesh(nodes, ele, MU, LAM, device=gpu, halfspace=True)
    xs = np.linspace(-3e3, 3e3, 9)
    X, Y = np.meshgrid(xs, xs + 125.0)      # no point on the fault trace y = 0
    obs = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], 1)
    got = dm.displacement(obs, 0.0, 1.0, nu=0.25).cpu().numpy()
    e = ele - 1
    ref = np.zeros_like(got)
    for j in range(dm.n):
        u = tde.TDdispHS(obs[:, 0], obs[:, 1], obs[:, 2], nodes[e[j, 0]], nodes[e[j, 1]],
                         nodes[e[j, 2]], 0.0, 1.0, 0.0, 0.25)
        ref += np.stack(u, 1)
    assert _close(got, ref, float(dm.n)).all()


def test_empty_and_invalid(gpu, pq3d_lib):
    out = torch.full((5, 3), 7.0, dtype=torch.float64, device=gpu)
    obs = torch.zeros((5, 3), dtype=torch.float64, device=gpu)
    # no sources: the sum is zero
    _lib.call("pq3d_tde_disp_sum", _lib.ptr(obs), 5, None, None, 0, None, None, None, NU,
              _lib.ptr(out), _lib.stream_handle())
    torch.cuda.synchronize()
    assert (out == 0).all()
    # no points: nothing written
    _lib.call("pq3d_tde_disp_sum", _lib.ptr(obs), 0, None, None, 0, None, None, None, NU,
              _lib.ptr(out), _lib.stream_handle())
    with pytest.raises(RuntimeError):
        _lib.call("pq3d_tde_disp_sum", _lib.ptr(obs), 5, None, None, 3, None, None, None, NU,
                  _lib.ptr(out), _lib.stream_handle())
    with pytest.raises(RuntimeError):
        _lib.call("pq3d_tde_disp_points", _lib.ptr(obs), _lib.ptr(obs), _lib.ptr(obs), 5, None,
                  None, 1.0, 0.0, 0.0, NU, _lib.ptr(out), _lib.stream_handle())
    m = planar_fault(length=4e3, depth=2e3, nx=2, nz=2)
    dm = tde.DeviceMesh(m.nodelst, m.elelst, MU, LAM, device=gpu)
    with pytest.raises(ValueError):
        dm.displacement(np.zeros((3, 2)), 1.0, 0.0)
    with pytest.raises(ValueError):
        dm.displacement(np.zeros((3, 3)), np.ones(dm.n + 1), 0.0)
