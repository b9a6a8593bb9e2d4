"""The no-allocation H-matvec entry point (pq3d_hmat_matvec_noalloc) against
the allocating one (pq3d_hmat_matvec) through the C-ABI: bitwise equal
products on two calls in a row, the second on a smaller operator (the static
scratch keeps the first call's contents), and a product larger than the
scratch rejected with a clear error."""
import numpy as np
import pytest
import torch

from pyquake3d_amd import _lib
from pyquake3d_amd.hmatrix import HMatrix
from pyquake3d_amd.mesh import undulating_fault

pytestmark = pytest.mark.gpu

MU = LAM = 32e9
SCRATCH_DOUBLES = 8 << 20


def _noalloc(H, v1, v2, out, nrhs, nt=None, npart=None):
    red = H._bufs(nrhs)[2]
    _lib.call("pq3d_hmat_matvec_noalloc", _lib.ptr(H.units), H.nunits, _lib.ptr(H.members),
              _lib.ptr(red), int(red.shape[0]), _lib.ptr(H.tiles), _lib.ptr(H.tile_ptr),
              H.ntiles, _lib.ptr(H.items), _lib.ptr(H.pool), H.esize, nrhs, _lib.ptr(v1),
              _lib.ptr(v2), H.nt if nt is None else nt, H.npart if npart is None else npart,
              _lib.ptr(out[0]), _lib.ptr(out[1]), _lib.ptr(out[2]),
              H.n if nrhs > 1 else 0, _lib.stream_handle())


def test_noalloc_matvec_equals_allocating_matvec(gpu, pq3d_lib):
    rng = np.random.default_rng(11)
    # the second call runs on a smaller operator than the first; the first
    # has blocks split over several phase-1 units (partial rows in use)
    for (L, D, nx, nz), leaf, dtype, nrhs in (((100e3, 25e3, 112, 28), 64, torch.float32, 4),
                                             ((16e3, 8e3, 16, 8), 16, torch.float64, 1)):
        mesh = undulating_fault(length=L, depth=D, nx=nx, nz=nz, seed=2)
        H = HMatrix.from_mesh(mesh, MU, LAM, leaf=leaf, eta=2.0, eps=1e-6, dtype=dtype,
                              device=gpu)
        assert H.stats.n_lowrank > 0
        assert H.nred > 0 or leaf == 16
        assert (H.nt + H.npart) * nrhs <= SCRATCH_DOUBLES
        n = H.n
        v1 = torch.from_numpy(rng.standard_normal(n * nrhs)).to(gpu)
        v2 = torch.from_numpy(rng.standard_normal(n * nrhs)).to(gpu)
        ref = H.matvec(v1, v2, nrhs=nrhs, instance_major=nrhs > 1)
        out = torch.full((3, n * nrhs), float("nan"), dtype=torch.float64, device=gpu)
        _noalloc(H, v1, v2, out, nrhs)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (nx, nz)


def test_noalloc_matvec_rejects_products_beyond_its_scratch(gpu, pq3d_lib):
    mesh = undulating_fault(length=16e3, depth=8e3, nx=16, nz=8, seed=2)
    H = HMatrix.from_mesh(mesh, MU, LAM, leaf=16, eta=2.0, eps=1e-6, dtype=torch.float32,
                          device=gpu)
    n = H.n
    v = torch.zeros(n * 16, dtype=torch.float64, device=gpu)
    out = torch.zeros((3, n * 16), dtype=torch.float64, device=gpu)
    with pytest.raises(RuntimeError, match="at most 8388608"):
        _noalloc(H, v, v, out, 16, nt=SCRATCH_DOUBLES // 16, npart=1)
    assert bool((out == 0).all())
