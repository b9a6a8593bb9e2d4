"""Device construction of the cluster tree and block partition.

CPU: the level-by-level drivers (node numbering, child lists, path keys and
their sort into depth-first order) run against a numpy stand-in of one level
of each kernel and must give the host builders' tree and partition.
GPU: the HIP kernels give them bit for bit (permutation, ranges, boxes,
block sequence and admissibility)."""
import numpy as np
import pytest
import torch

from pyquake3d_amd import cluster
from pyquake3d_amd.cluster import (build_block_partition, build_block_partition_device,
                                   build_cluster_tree, build_cluster_tree_device)
from pyquake3d_amd.mesh import get_eleVec, planar_fault, undulating_fault

from cluster_emulation import fake_block_level as _fake_block_level
from cluster_emulation import fake_tree_level as _fake_tree_level


def _clouds():
    out = []
    for mesh in (planar_fault(nx=12, nz=6), undulating_fault(nx=40, nz=10)):
        out.append(get_eleVec(mesh.nodelst, mesh.elelst)[1])
    rng = np.random.default_rng(3)
    pts = rng.normal(size=(700, 3)) * [5e3, 1e3, 2e3]
    pts[100:160] = pts[100]                     # a cluster of equal points
    pts[300:340, 1] = 7.0                       # equal coordinates on one axis
    out.append(pts)
    return out


def _same_tree(td, th):
    assert np.array_equal(td.perm, th.perm)
    key = lambda t: {(int(s), int(m)): (tuple(lo), tuple(hi), t.left[k] < 0)
                     for k, (s, m, lo, hi) in enumerate(zip(t.start, t.size, t.lo, t.hi))}
    kd, kh = key(td), key(th)
    assert kd == kh and td.n_nodes == th.n_nodes
    assert len(td.leaves()) == len(th.leaves())
    assert np.array_equal(td.start[td.leaves()], th.start[th.leaves()])


def _same_partition(Pd, td, Ph, th):
    seq = lambda P, t: np.stack([t.start[P.rows], t.size[P.rows], t.start[P.cols],
                                 t.size[P.cols], P.adm.astype(np.int64)], 1)
    assert np.array_equal(seq(Pd, td), seq(Ph, th))


@pytest.mark.parametrize("leaf,eta,min_size", [(16, 2.0, None), (32, 2.0, None),
                                               (64, 1.0, 8), (8, 3.0, 32)])
def test_driver_matches_host_builders(monkeypatch, leaf, eta, min_size):
    monkeypatch.setattr(cluster, "_tree_level", _fake_tree_level)
    monkeypatch.setattr(cluster, "_block_level", _fake_block_level)
    for pts in _clouds():
        th = build_cluster_tree(pts, leaf)
        td = build_cluster_tree_device(pts, leaf, device="cpu")
        _same_tree(td, th)
        Ph = build_block_partition(th, eta, min_size)
        _same_partition(build_block_partition_device(td, eta, min_size, device="cpu"), td, Ph, th)
        # host tree numbering through the device partition gives the host block list
        Pd = build_block_partition_device(th, eta, min_size, device="cpu")
        assert np.array_equal(Pd.rows, Ph.rows) and np.array_equal(Pd.cols, Ph.cols)
        assert np.array_equal(Pd.adm, Ph.adm)


@pytest.mark.gpu
@pytest.mark.parametrize("leaf,eta,min_size", [(16, 2.0, None), (64, 2.0, None), (8, 3.0, 32)])
def test_device_kernels_match_host_builders(gpu, pq3d_lib, leaf, eta, min_size):
    for pts in _clouds():
        th = build_cluster_tree(pts, leaf)
        td = build_cluster_tree_device(pts, leaf, device=gpu)
        _same_tree(td, th)
        Ph = build_block_partition(th, eta, min_size)
        _same_partition(build_block_partition_device(td, eta, min_size, device=gpu), td, Ph, th)
        Pd = build_block_partition_device(th, eta, min_size, device=gpu)
        assert np.array_equal(Pd.rows, Ph.rows) and np.array_equal(Pd.cols, Ph.cols)
        assert np.array_equal(Pd.adm, Ph.adm)


@pytest.mark.gpu
def test_device_tree_full_size(gpu, pq3d_lib):
    """N = 100,352 (the undulating 448 x 112 fault), leaf 64, eta 2."""
    mesh = undulating_fault(nx=448, nz=112)
    xg = get_eleVec(mesh.nodelst, mesh.elelst)[1]
    th = build_cluster_tree(xg, 64)
    td = build_cluster_tree_device(xg, 64, device=gpu)
    _same_tree(td, th)
    _same_partition(build_block_partition_device(td, 2.0, device=gpu), td,
                    build_block_partition(th, 2.0), th)
