"""Step-size MLP kernels (forward, backward, Adam) and the feature reduction
against the numpy oracle of the same network and numpy reductions."""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle
from pyquake3d_amd import _lib

pytestmark = pytest.mark.gpu


def _mlp(gpu, seed=0):
    from pyquake3d_amd.learned_dt import StepMLP
    return StepMLP(device=gpu, seed=seed)


def test_parameter_count_matches_oracle(gpu, pq3d_lib):
    assert _mlp(gpu).P == mlp_oracle.P == 8833


@pytest.mark.parametrize("nsamp,chunk", [(1, 64), (300, 64), (257, 16)])
def test_forward_backward_match_oracle(gpu, pq3d_lib, nsamp, chunk):
    """FP32 kernels vs FP64 numpy: loss and every gradient entry within 1e-4
    of the gradient's largest magnitude (ragged last chunk included)."""
    net = _mlp(gpu, seed=1)
    rng = np.random.default_rng(nsamp)
    X = rng.standard_normal((nsamp, 6)).astype(np.float32)
    T = rng.standard_normal(nsamp).astype(np.float32)
    p = net.prm.cpu().numpy()
    # non-zero biases so that their gradients are exercised through tanh'
    p = p + rng.normal(0, 0.1, p.shape).astype(np.float32)
    net.prm.copy_(torch.from_numpy(p))
    Xd, Td = torch.from_numpy(X).to(gpu), torch.from_numpy(T).to(gpu)
    loss, g = net.loss_and_grad(Xd, Td, chunk=chunk)
    ref_loss, ref_g = mlp_oracle.loss_and_grad(p.astype(np.float64), X, T.astype(np.float64))
    assert loss == pytest.approx(ref_loss, rel=1e-4)
    g = g.cpu().numpy().astype(np.float64)
    assert np.abs(g - ref_g).max() <= 1e-4 * np.abs(ref_g).max()
    y = net.predict_raw(Xd).cpu().numpy()
    ref_y, _ = mlp_oracle.forward(p, X.astype(np.float64))
    assert np.abs(y - ref_y).max() <= 1e-5 * max(1.0, np.abs(ref_y).max())


def test_adam_step_matches_numpy(gpu, pq3d_lib):
    """Three Adam steps: the kernel's update from its own summed gradient
    equals numpy's Adam applied to that gradient."""
    net = _mlp(gpu, seed=2)
    rng = np.random.default_rng(5)
    X = rng.standard_normal((200, 6)).astype(np.float32)
    T = rng.standard_normal(200).astype(np.float32)
    m = np.zeros(net.P)
    v = np.zeros(net.P)
    lr, b1, b2, eps = 1e-2, 0.9, 0.999, 1e-8
    Xd, Td = torch.from_numpy(X).to(gpu), torch.from_numpy(T).to(gpu)
    for t in range(1, 4):
        p = net.prm.cpu().numpy().astype(np.float64)
        _, gd = net.loss_and_grad(Xd, Td)  # lr = 0: gradient only, moments untouched
        g = gd.cpu().numpy().astype(np.float64)
        grad, _, nblk = net._fb(Xd, Td)
        _lib.call("pq3d_mlp_adam", _lib.ptr(net.prm), _lib.ptr(grad), nblk, 200,
                  _lib.ptr(net.m), _lib.ptr(net.v), lr, b1, b2, eps, t, None,
                  _lib.stream_handle())
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
        got = net.prm.cpu().numpy().astype(np.float64)
        assert np.abs(got - p).max() <= 1e-3 * lr, t


def test_training_reduces_loss(gpu, pq3d_lib):
    """Fit a smooth function of the six inputs; the kernels' Adam loop must
    reduce the normalised loss by an order of magnitude."""
    net = _mlp(gpu, seed=3)
    rng = np.random.default_rng(7)
    X = rng.uniform(-1, 1, (4096, 6)).astype(np.float32)
    T = (np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2] - 0.5 * X[:, 3] ** 2).astype(np.float32)
    hist = net.fit(X, T, epochs=40, batch=512, lr=3e-3)
    assert hist[-1] < 0.1 * hist[0]
    pred = net.predict_log_dt(X)
    assert np.sqrt(np.mean((pred - T) ** 2)) < 0.3 * T.std()


@pytest.mark.parametrize("ragged", [False, True])
def test_features_match_numpy(gpu, pq3d_lib, ragged):
    """Per-instance reductions over instance-major cells (equal or uneven
    instance sizes) against numpy."""
    from pyquake3d_amd.learned_dt import decode_features
    rng = np.random.default_rng(11)
    nb = 7
    sizes = rng.integers(1, 6000, nb) if ragged else np.full(nb, 5003)
    inst = np.repeat(np.arange(nb, dtype=np.int32), sizes)
    ne = int(sizes.sum())
    V = 10.0 ** rng.uniform(-14, 0.5, ne)
    psi = rng.uniform(0.5, 0.8, ne)
    b = rng.uniform(0.01, 0.02, ne)
    f0 = np.full(ne, 0.6)
    V0 = 1e-6
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(gpu)
    acc = torch.zeros(4 * nb, dtype=torch.int64, device=gpu)
    dI, dV, dpsi, db, df0 = d(inst), d(V), d(psi), d(b), d(f0)  # held until the kernel has run
    _lib.call("pq3d_features", ne, nb, _lib.ptr(dI), _lib.ptr(dV), _lib.ptr(dpsi), _lib.ptr(db),
              _lib.ptr(df0), V0, _lib.ptr(acc), _lib.stream_handle())
    vmax, vmin, mean, frac = decode_features(acc.cpu().numpy().reshape(nb, 4), sizes)
    lv = np.log10(V)
    lvt = np.log10(V / V0) + (psi - f0) / (b * np.log(10.0))
    for k in range(nb):
        s = inst == k
        # 1000 + log10 V is stored as a double: resolution 2^-43 * 1000
        assert abs(vmax[k] - lv[s].max()) <= 1e-12
        assert abs(vmin[k] - lv[s].min()) <= 1e-12
        assert abs(mean[k] - lvt[s].mean()) <= 1e-8
        assert frac[k] == (V[s] > 1e-3).sum() / sizes[k]


def test_learned_step_size_loop(gpu, pq3d_lib):
    """Learned first-attempt step size on the 144-cell problem: trained with
    the kernels on (features, log10 h_accepted/dtnext) pairs logged from the
    reference controller, it cuts the rejected attempts per accepted step;
    the embedded error test still decides every step (no accepted step has
    max(c1, c2) >= 1) and the PI fallback handles the rejections. The
    operator product is frozen per step, so a rejected attempt costs one
    cell-local RK45 attempt, not an operator product: this is a gain in
    attempts, not in operator products per simulated year."""
    from pyquake3d_amd.learned_dt import FeatureEnsemble, StepMLP
    from pyquake3d_amd.problem import planar_c1
    from pyquake3d_amd.qdsim import DenseOperator
    p = planar_c1(nx=12, nz=6, seed=0)
    op = DenseOperator(p.mesh, p.mu, p.lambda_, device=gpu)
    rec = FeatureEnsemble([p], op, device=gpu)
    rec.record = True
    for _ in range(1500):
        rec.step()
    plain_aps = rec.nattempts / rec.nsteps
    X, T = rec.dataset()
    assert (T <= 0).all() and (T < 0).any()
    mlp = StepMLP(device=gpu, seed=0)
    hist = mlp.fit(X, T, epochs=60, batch=1024, lr=2e-3)
    assert hist[-1] < hist[0]
    le = FeatureEnsemble([p], op, mlp=mlp, device=gpu)
    for _ in range(1500):
        le.step()
    learned_aps = le.nattempts / le.nsteps
    assert learned_aps < plain_aps, (learned_aps, plain_aps)
    assert le.max_accepted_ratio < 1.0
    assert le.nfallback > 0
