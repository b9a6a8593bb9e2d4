"""GPU H-matrix: near-field assembly, in-kernel ACA and the streamed matvec,
against the dense FP64 product (CPU oracle operators)."""
import numpy as np
import pytest
import torch

from pyquake3d_amd.hmatrix import (MFP_DEPTHS, MFP_WAVES, MFS_DEFAULT, MFS_DEPTHS,
                                   MFS_ORDERS, MFS_WAVES, HMatrix)
from pyquake3d_amd.mesh import get_eleVec, undulating_fault

pytestmark = pytest.mark.gpu
MU = LAM = 32e9


@pytest.fixture(scope="module")
def case(oracle):
    mesh = undulating_fault(length=40e3, depth=12e3, nx=44, nz=18, seed=4)
    ev, xg = get_eleVec(mesh.nodelst, mesh.elelst)
    n = mesh.n
    K = oracle.traction_block(xg, ev, mesh.triangles(), np.arange(n), np.arange(n), MU, LAM)
    return mesh, K


def _apply(H, v1, v2):
    dev = H.device
    x1 = torch.from_numpy(v1[H.perm]).to(dev)
    x2 = torch.from_numpy(v2[H.perm]).to(dev)
    y = H.matvec(x1, x2).cpu().numpy()
    out = np.empty_like(y)
    out[:, H.perm] = y
    return out


def _dense(K, v1, v2):
    return np.stack([K[0] @ v1 + K[3] @ v2, K[1] @ v1 + K[4] @ v2, K[2] @ v1 + K[5] @ v2])


def _relerr(y, ref):
    return np.abs(y - ref).max() / np.abs(ref).max()


def test_all_dense_hmatrix_equals_dense_product(gpu, pq3d_lib, case):
    mesh, K = case
    H = HMatrix.from_mesh(mesh, MU, LAM, leaf=32, eta=0.0, dtype=torch.float64, device=gpu)
    assert H.stats.n_lowrank == 0
    rng = np.random.default_rng(1)
    v1, v2 = rng.standard_normal(mesh.n), rng.standard_normal(mesh.n)
    assert _relerr(_apply(H, v1, v2), _dense(K, v1, v2)) < 1e-12


def test_fp64_aca_matvec_tight(gpu, pq3d_lib, case):
    """eps=1e-7: blocks that ACA cannot compress below their dense size (or
    that reach the rank cap) are stored densely; the product stays within
    1e-5 of the dense FP64 product."""
    mesh, K = case
    H = HMatrix.from_mesh(mesh, MU, LAM, leaf=32, eta=2.0, eps=1e-7, dtype=torch.float64,
                          device=gpu)
    assert H.stats.n_lowrank > 0
    rng = np.random.default_rng(2)
    for _ in range(3):
        v1, v2 = rng.standard_normal(mesh.n), rng.standard_normal(mesh.n)
        assert _relerr(_apply(H, v1, v2), _dense(K, v1, v2)) < 1e-5


def test_aca_blocks_meet_relative_frobenius_bound(gpu, pq3d_lib, case):
    mesh, K = case
    eps = 1e-4
    H = HMatrix.from_mesh(mesh, MU, LAM, leaf=32, eta=2.0, eps=eps, dtype=torch.float64,
                          device=gpu)
    B = H.blocks
    perm = H.perm
    worst = 0.0
    for q, bi in enumerate(B["lr_idx"]):
        r = perm[B["r0"][bi]:B["r0"][bi] + B["m"][bi]]
        c = perm[B["c0"][bi]:B["c0"][bi] + B["n"][bi]]
        S = np.block([[K[0][np.ix_(r, c)], K[3][np.ix_(r, c)]],
                      [K[1][np.ix_(r, c)], K[4][np.ix_(r, c)]],
                      [K[2][np.ix_(r, c)], K[5][np.ix_(r, c)]]])
        U, V = H.lowrank_factors(q)
        worst = max(worst, np.linalg.norm(S - U @ V) / np.linalg.norm(S))
    # ACA's stopping test estimates the next term; allow a factor 10 over eps
    assert worst < 10 * eps, worst


def test_fp32_matvec_within_stated_bound(gpu, pq3d_lib, case):
    """FP32 storage + ACA eps=1e-4: max-norm relative error of the three
    traction rates <= 1e-3 (stated bound); FP32 rounding alone is ~1e-7."""
    mesh, K = case
    H = HMatrix.from_mesh(mesh, MU, LAM, leaf=32, eta=2.0, eps=1e-4, dtype=torch.float32,
                          device=gpu)
    H64 = HMatrix.from_mesh(mesh, MU, LAM, leaf=32, eta=2.0, eps=1e-4, dtype=torch.float64,
                            device=gpu)
    rng = np.random.default_rng(3)
    v1, v2 = rng.standard_normal(mesh.n), rng.standard_normal(mesh.n)
    ref = _dense(K, v1, v2)
    y32 = _apply(H, v1, v2)
    y64 = _apply(H64, v1, v2)
    assert _relerr(y32, ref) < 1e-3
    # the FP32 storage adds only rounding on top of the same ACA factors
    assert _relerr(y32, y64) < 1e-5
    # a deficit-like vector: smooth slip-rate pattern
    xg = get_eleVec(mesh.nodelst, mesh.elelst)[1]
    v1 = np.exp(-((xg[:, 0] / 8e3) ** 2 + ((xg[:, 2] + 6e3) / 3e3) ** 2)) * 1e-6
    v2 = 0.1 * v1
    assert _relerr(_apply(H, v1, v2), _dense(K, v1, v2)) < 1e-3


@pytest.mark.parametrize("dtype,leaf", [(torch.float64, 32), (torch.float32, 64)])
def test_multi_rhs_products_match_single(gpu, pq3d_lib, case, dtype, leaf):
    """Batched products (NB = 2, 4, 8, 16) equal the single-vector product bit
    for bit in every column, with interleaved and with instance-major
    outputs. Ranks above 64 exercise the second pass over rank chunks."""
    mesh, _ = case
    H = HMatrix.from_mesh(mesh, MU, LAM, leaf=leaf, eta=2.0, eps=1e-9, dtype=dtype, device=gpu,
                          kmax=128)
    n = H.n
    rng = np.random.default_rng(7)
    X1 = torch.from_numpy(rng.standard_normal((n, 16))).to(gpu)
    X2 = torch.from_numpy(rng.standard_normal((n, 16))).to(gpu)
    single = [H.matvec(X1[:, b].contiguous(), X2[:, b].contiguous()).clone() for b in range(16)]
    for nb in (2, 4, 8, 16):
        y = H.matvec(X1[:, :nb].contiguous().view(-1), X2[:, :nb].contiguous().view(-1),
                     nrhs=nb).view(3, n, nb)
        for b in range(nb):
            assert torch.equal(y[:, :, b], single[b]), (nb, b)
        yi = H.matvec(X1[:, :nb].contiguous().view(-1), X2[:, :nb].contiguous().view(-1),
                      nrhs=nb, instance_major=True).view(3, nb, n)
        for b in range(nb):
            assert torch.equal(yi[:, b, :], single[b]), ("instance-major", nb, b)


@pytest.mark.parametrize("mode", ["mfma", "mfma_comp"])
@pytest.mark.parametrize("leaf", [32, 64])
def test_mfma_matvec_16_rhs(gpu, pq3d_lib, case, leaf, mode):
    """FP32 storage, 16 right-hand sides, dense and U products on FP32 matrix
    cores: within the FP32 path's stated bound (1e-3 of the largest value)
    of the dense FP64 product, within 1e-5 of the FP64-accumulated product of
    the same storage (FP32 rounding of the weights and FP32 sums), identical
    on repeat, and a permutation of the right-hand sides permutes the
    outputs bit for bit (every column takes the same sequence of
    operations). Leaf 32 tiles use two 16-row blocks, leaf 64 four."""
    mesh, K = case
    H = HMatrix.from_mesh(mesh, MU, LAM, leaf=leaf, eta=2.0, eps=1e-4, dtype=torch.float32,
                          device=gpu)
    assert H.stats.n_lowrank > 0
    n = H.n
    rng = np.random.default_rng(17)
    V1, V2 = rng.standard_normal((n, 16)), rng.standard_normal((n, 16))
    V1[:, 3] = np.exp(-((H.xg[:, 0] / 8e3) ** 2)) * 1e-6       # a smooth deficit-like column
    V2[:, 3] = 0.1 * V1[:, 3]
    X1 = torch.from_numpy(np.ascontiguousarray(V1[H.perm])).to(gpu)
    X2 = torch.from_numpy(np.ascontiguousarray(V2[H.perm])).to(gpu)
    y64 = H.matvec(X1.view(-1), X2.view(-1), nrhs=16).view(3, n, 16).cpu().numpy()
    H.accumulate = mode
    y = H.matvec(X1.view(-1), X2.view(-1), nrhs=16).view(3, n, 16)
    assert torch.equal(y, H.matvec(X1.view(-1), X2.view(-1), nrhs=16).view(3, n, 16))
    yi = H.matvec(X1.view(-1), X2.view(-1), nrhs=16, instance_major=True).view(3, 16, n)
    assert torch.equal(yi.transpose(1, 2), y)
    y = y.cpu().numpy()
    for b in range(16):
        ref = _dense(K, V1[:, b], V2[:, b])
        got = np.empty_like(y[:, :, b])
        got[:, H.perm] = y[:, :, b]
        assert _relerr(got, ref) < 1e-3, b
        assert _relerr(y[:, :, b], y64[:, :, b]) < 1e-5, b
    p = rng.permutation(16)
    yp = H.matvec(X1[:, p].contiguous().view(-1), X2[:, p].contiguous().view(-1),
                  nrhs=16).view(3, n, 16).cpu().numpy()
    assert np.array_equal(yp, y[:, :, p])
    # other sizes and FP64 storage keep the FP64-accumulating kernels
    y1 = H.matvec(X1[:, 0].contiguous(), X2[:, 0].contiguous()).cpu().numpy()
    assert np.array_equal(y1, y64[:, :, 0])


@pytest.mark.parametrize("planes", [2, 3])
def test_step_stream_configurations_and_autotuner(gpu, pq3d_lib, case, planes, tmp_path):
    """Every configuration of the step-stream product (phase 2: steps in
    flight, wavefronts per tile, tile order; phase 1 over the V panels:
    steps in flight, wavefronts per unit) stays within 1e-5 of the
    FP64-accumulated product of the same storage and is repeatable; the
    autotuner times its candidates, keeps the fastest, and the tuned choice
    survives save/load."""
    mesh, _ = case
    H = HMatrix.from_mesh(mesh, MU, LAM, leaf=32, eta=2.0, eps=1e-4, dtype=torch.float32,
                          device=gpu)
    n = H.n
    rng = np.random.default_rng(23)
    X1 = torch.from_numpy(rng.standard_normal(n * 16)).to(gpu)
    X2 = torch.from_numpy(rng.standard_normal(n * 16)).to(gpu)
    y64 = H.matvec(X1, X2, nrhs=16, planes=planes).view(3, n, 16)[:planes].cpu().numpy()
    H.accumulate = "mfma"
    cfgs = [dict(MFS_DEFAULT, order=o, waves=w, depth=d)
            for o in MFS_ORDERS for w in MFS_WAVES for d in MFS_DEPTHS]
    cfgs += [dict(MFS_DEFAULT, p1_depth=d, p1_waves=w) for d in (2, 4) for w in (2, 4)]
    for cfg in cfgs:
        H.tuning = cfg
        y = H.matvec(X1, X2, nrhs=16, planes=planes).view(3, n, 16)[:planes]
        assert torch.equal(y, H.matvec(X1, X2, nrhs=16, planes=planes).view(3, n, 16)[:planes])
        y = y.cpu().numpy()
        for b in range(16):
            assert _relerr(y[:, :, b], y64[:, :, b]) < 1e-5, (cfg, b)
    H.tuning = None
    rep = H.tune(planes=planes, reps=3)
    assert len(rep["candidates"]) == (len(MFS_DEPTHS) * len(MFS_WAVES) * len(MFS_ORDERS)
                                      + len(MFP_DEPTHS) * len(MFP_WAVES) - 1)
    assert rep["choice"] == min(rep["candidates"], key=lambda c: c[1])[0]
    assert H.tuning == rep["choice"] and H.accumulate == "mfma"
    y = H.matvec(X1, X2, nrhs=16, planes=planes).view(3, n, 16)[:planes].cpu().numpy()
    for b in range(16):
        assert _relerr(y[:, :, b], y64[:, :, b]) < 1e-5
    H.save(tmp_path / "h.safetensors")
    G = HMatrix.load(tmp_path / "h.safetensors", device=gpu)
    assert G.tuning == H.tuning
    G.accumulate = "mfma"
    yg = G.matvec(X1, X2, nrhs=16, planes=planes).view(3, n, 16)[:planes].cpu().numpy()
    assert np.array_equal(yg, y)


def test_halfspace_hmatrix_matches_dense_halfspace_operators(gpu, pq3d_lib):
    """Half-space H-matrix of a surface-breaking fault: all-dense storage
    equals the half-space traction block; ACA at eps 1e-7 stays within 1e-5
    of the dense product."""
    from pyquake3d_amd import tde
    from pyquake3d_amd.mesh import planar_fault
    mesh = planar_fault(length=16e3, depth=8e3, nx=16, nz=8, seed=3)
    dm = tde.DeviceMesh.from_mesh(mesh, MU, LAM, device=gpu, halfspace=True)
    K = dm.traction_block().cpu().numpy()
    rng = np.random.default_rng(4)
    v1, v2 = rng.standard_normal(mesh.n), rng.standard_normal(mesh.n)
    ref = _dense(K, v1, v2)
    Hd = HMatrix.from_mesh(mesh, MU, LAM, leaf=16, eta=0.0, dtype=torch.float64, device=gpu,
                           halfspace=True)
    assert _relerr(_apply(Hd, v1, v2), ref) < 1e-12
    Ha = HMatrix.from_mesh(mesh, MU, LAM, leaf=16, eta=2.0, eps=1e-7, dtype=torch.float64,
                           device=gpu, halfspace=True)
    assert Ha.stats.n_lowrank > 0
    assert _relerr(_apply(Ha, v1, v2), ref) < 1e-5
    # the free surface matters: the full-space operators differ
    K_fs = tde.DeviceMesh.from_mesh(mesh, MU, LAM, device=gpu).traction_block().cpu().numpy()
    assert np.abs(K_fs - K).max() > 1e-3 * np.abs(K).max()


def test_full_size_fp32_hmatrix_sampled_rows(gpu, pq3d_lib):
    """N = 100,352 undulating fault, leaf 64, eta 2, ACA eps 1e-4, FP32
    storage: 256 sampled rows of the product against the FP64 dense rows
    (evaluated on the device, all 100,352 columns) for a random and a smooth
    deficit-like vector: max error <= 1e-3 of the largest value."""
    from pyquake3d_amd import tde
    from pyquake3d_amd.problem import undulating_c3
    p = undulating_c3()
    H = HMatrix.from_mesh(p.mesh, MU, LAM, leaf=64, eta=2.0, eps=1e-4, dtype=torch.float32,
                          device=gpu)
    dm = tde.DeviceMesh.from_mesh(p.mesh, MU, LAM, device=gpu)
    n = H.n
    rng = np.random.default_rng(11)
    rows = np.sort(rng.choice(n, 256, replace=False))          # original element order
    Kr = dm.traction_block(rows, None)                          # (6, 256, n) FP64
    xg = get_eleVec(p.mesh.nodelst, p.mesh.elelst)[1]
    smooth = np.exp(-((xg[:, 0] / 20e3) ** 2 + ((xg[:, 2] + 10e3) / 6e3) ** 2)) * 1e-6
    for v1, v2 in ((rng.standard_normal(n), rng.standard_normal(n)), (smooth, 0.2 * smooth)):
        x1 = torch.from_numpy(v1).to(gpu)
        x2 = torch.from_numpy(v2).to(gpu)
        ref = torch.stack([Kr[c] @ x1 + Kr[3 + c] @ x2 for c in range(3)]).cpu().numpy()
        y = _apply(H, v1, v2)[:, rows]
        assert _relerr(y, ref) < 1e-3
    # 16 right-hand sides on FP32 matrix cores: scaled copies of the two vectors
    H.accumulate = "mfma"
    v = [rng.standard_normal(n), smooth]
    sc = 10.0 ** rng.uniform(-3, 3, 16)
    V1 = np.stack([sc[b] * v[b % 2] for b in range(16)], axis=1)
    V2 = np.stack([0.5 * sc[b] * v[(b + 1) % 2] for b in range(16)], axis=1)
    X1 = torch.from_numpy(np.ascontiguousarray(V1[H.perm])).to(gpu)
    X2 = torch.from_numpy(np.ascontiguousarray(V2[H.perm])).to(gpu)
    Y = H.matvec(X1.view(-1), X2.view(-1), nrhs=16, instance_major=True).view(3, 16, n)
    Y = Y[:, :, torch.from_numpy(H.iperm).to(gpu)][:, :, torch.from_numpy(rows).to(gpu)]
    for b in range(16):
        x1 = torch.from_numpy(V1[:, b]).to(gpu)
        x2 = torch.from_numpy(V2[:, b]).to(gpu)
        ref = torch.stack([Kr[c] @ x1 + Kr[3 + c] @ x2 for c in range(3)]).cpu().numpy()
        assert _relerr(Y[:, b].cpu().numpy(), ref) < 1e-3, b


def test_saved_hmatrix_reloads_bit_for_bit(gpu, pq3d_lib, case, tmp_path):
    """An H-matrix written with save() and read with load() gives the same
    products (1 and 8 right-hand sides) bit for bit, and its low-rank
    factors are the same."""
    mesh, _ = case
    H = HMatrix.from_mesh(mesh, MU, LAM, leaf=32, eta=2.0, eps=1e-6, dtype=torch.float32,
                          device=gpu)
    f = tmp_path / "h.safetensors"
    H.save(f, leaf="32")
    G = HMatrix.load(f, device=gpu)
    assert G.meta == {"leaf": "32"} and G.stats.n_lowrank == H.stats.n_lowrank
    n = H.n
    rng = np.random.default_rng(5)
    for nb in (1, 8):
        v1 = torch.from_numpy(rng.standard_normal(n * nb)).to(gpu)
        v2 = torch.from_numpy(rng.standard_normal(n * nb)).to(gpu)
        assert torch.equal(H.matvec(v1, v2, nrhs=nb), G.matvec(v1, v2, nrhs=nb)), nb
    if H.stats.n_lowrank:
        for a, b in zip(H.lowrank_factors(0), G.lowrank_factors(0)):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("halfspace", [False, True])
def test_wide_aca_equals_one_workgroup_aca(gpu, pq3d_lib, case, halfspace):
    """Every admissible block through the wide ACA (rank steps spread over
    many workgroups) against every block through the one-workgroup ACA: the
    same ranks and the same stored factors bit for bit, full and half space,
    FP64 and FP32 storage (with a rank cap of 12 some blocks reach the cap
    and fall back to dense storage)."""
    mesh, _ = case
    for dtype, eps, kmax in ((torch.float64, 1e-7, 128), (torch.float32, 1e-4, 12)):
        kw = dict(leaf=32, eta=2.0, eps=eps, dtype=dtype, device=gpu, halfspace=halfspace,
                  kmax=kmax)
        H1 = HMatrix.from_mesh(mesh, MU, LAM, aca_wide=1 << 62, **kw)
        Hw = HMatrix.from_mesh(mesh, MU, LAM, aca_wide=0, **kw)
        if kmax == 128:
            assert H1.stats.n_lowrank > 0
        else:
            assert H1.stats.fallback_dense > 0
        assert np.array_equal(H1.blocks["ranks"], Hw.blocks["ranks"])
        assert np.array_equal(H1.blocks["dense_idx"], Hw.blocks["dense_idx"])
        assert torch.equal(H1.pool, Hw.pool)


def test_wide_aca_at_config_size(gpu, pq3d_lib):
    """N = 100,352 (leaf 64, eta 2, eps 1e-4): the default split (blocks with
    m + n >= 4,096 wide) stores the same operator as the one-workgroup ACA."""
    from pyquake3d_amd.problem import undulating_c3
    p = undulating_c3()
    kw = dict(leaf=64, eta=2.0, eps=1e-4, dtype=torch.float32, device=gpu)
    Hd = HMatrix.from_mesh(p.mesh, p.mu, p.lambda_, **kw)
    pool, ranks, t_default = Hd.pool.cpu(), Hd.blocks["ranks"].copy(), Hd.stats.build_s["aca"]
    del Hd
    torch.cuda.empty_cache()
    H1 = HMatrix.from_mesh(p.mesh, p.mu, p.lambda_, aca_wide=1 << 62, **kw)
    print(f"ACA: default split {t_default:.3f} s, one workgroup per block "
          f"{H1.stats.build_s['aca']:.3f} s")
    assert np.array_equal(H1.blocks["ranks"], ranks)
    assert torch.equal(H1.pool.cpu(), pool)


def test_fp32_storage_error_per_block(gpu, pq3d_lib, case):
    """Per-block error statement of the FP32 storage against the FP64 dense
    oracle: every near-field block within FP32 rounding (relative Frobenius
    error <= 1e-6), every low-rank block within the ACA bound (10 eps, the
    same allowance as the FP64 factors) and at most FP32 rounding above the
    FP64 factors of the same block."""
    mesh, K = case
    eps = 1e-4
    H = HMatrix.from_mesh(mesh, MU, LAM, leaf=32, eta=2.0, eps=eps, dtype=torch.float32,
                          device=gpu)
    H64 = HMatrix.from_mesh(mesh, MU, LAM, leaf=32, eta=2.0, eps=eps, dtype=torch.float64,
                            device=gpu)
    B = H.blocks
    perm = H.perm
    assert np.array_equal(B["lr_idx"], H64.blocks["lr_idx"])

    def exact(bi):
        r = perm[B["r0"][bi]:B["r0"][bi] + B["m"][bi]]
        c = perm[B["c0"][bi]:B["c0"][bi] + B["n"][bi]]
        return [K[o][np.ix_(r, c)] for o in range(6)]

    worst_dense = 0.0
    for q, bi in enumerate(B["dense_idx"]):
        m, n, base = int(B["m"][bi]), int(B["n"][bi]), int(B["d_base"][q])
        St = H.pool[base:base + 6 * m * n].double().cpu().numpy().reshape(6, n, m)
        T = np.stack(exact(bi))
        worst_dense = max(worst_dense,
                          np.linalg.norm(St.transpose(0, 2, 1) - T) / np.linalg.norm(T))
    assert worst_dense < 1e-6, worst_dense
    worst_lr = worst_gap = 0.0
    for q, bi in enumerate(B["lr_idx"]):
        T = exact(bi)
        S = np.block([[T[0], T[3]], [T[1], T[4]], [T[2], T[5]]])
        U, V = H.lowrank_factors(q)
        U64, V64 = H64.lowrank_factors(q)
        e32 = np.linalg.norm(U @ V - S) / np.linalg.norm(S)
        e64 = np.linalg.norm(U64 @ V64 - S) / np.linalg.norm(S)
        worst_lr = max(worst_lr, e32)
        worst_gap = max(worst_gap, abs(e32 - e64))
    assert worst_lr < 10 * eps, worst_lr
    assert worst_gap < 1e-6, worst_gap


def test_prefilled_weights_equal_fp64_inputs(gpu, pq3d_lib, case):
    """pq3d_deficit_w32 writes the interleaved deficit of instance-major slip
    rates into the matrix-core weights as the FP32 rounding of the FP64
    differences; the product on those prefilled weights (v1 = v2 = None)
    equals the product on the FP64 deficit bit for bit."""
    from pyquake3d_amd import _lib
    mesh, _ = case
    H = HMatrix.from_mesh(mesh, MU, LAM, leaf=32, eta=2.0, eps=1e-4, dtype=torch.float32,
                          device=gpu)
    H.accumulate = "mfma"
    n, nb = H.n, 16
    rng = np.random.default_rng(31)
    sv1 = torch.from_numpy(rng.standard_normal(nb * n) * 1e-9).to(gpu)   # instance-major
    sv2 = torch.from_numpy(rng.standard_normal(nb * n) * 1e-9).to(gpu)
    pl1 = torch.from_numpy(rng.standard_normal(nb * n) * 1e-9).to(gpu)
    pl2 = torch.from_numpy(rng.standard_normal(nb * n) * 1e-9).to(gpu)
    x1 = (sv1 - pl1).view(nb, n).t().contiguous().view(-1)      # interleaved FP64 deficit
    x2 = (sv2 - pl2).view(nb, n).t().contiguous().view(-1)
    ref = H.matvec(x1, x2, nrhs=nb, instance_major=True, planes=2).clone()
    W = H.weights()
    W[:2 * n * nb].fill_(float("nan"))
    _lib.call("pq3d_deficit_w32", n, nb, _lib.ptr(sv1), _lib.ptr(sv2), _lib.ptr(pl1),
              _lib.ptr(pl2), _lib.ptr(W), _lib.stream_handle())
    assert torch.equal(W[:n * nb], x1.float()) and torch.equal(W[n * nb:2 * n * nb], x2.float())
    got = H.matvec(None, None, nrhs=nb, instance_major=True, planes=2)
    assert torch.equal(got[:2], ref[:2])
    with pytest.raises(ValueError):
        H.accumulate = "f64"
        H.matvec(None, None, nrhs=nb)
