"""Learned step-size proposal for the RK45 stepper.

A 6 -> 64 -> 64 -> 64 -> 1 MLP (FP32, tanh) maps the solver state of each
simulation to log10 of the next step size. Inputs per instance:
  [max log10 V, min log10 V, mean log10(V theta/Dc), log10 last dt,
   log10 last error ratio, fraction of cells with V > 1e-3 m/s]
(the cell reductions are one HIP kernel, ``pq3d_features``). The network is
trained on the GPU with the library's own forward/backward/Adam kernels on
(state, target) pairs logged from runs of the plain controller; the target is
log10 of the controller's own estimate of the largest acceptable step at the
accepted attempt, 0.9 h (tol/err)^0.2, without the reference's 1.5 h growth
cap (src/QDsim.py:941-942).

In ``LearnedEnsemble`` every step still goes through the reference's embedded
error test: the proposal is the first attempt's h; a rejected attempt falls
back to the PI controller (Gustafsson's h (tol/err)^0.14 (err_prev/tol)^0.08
with 0.9 safety), as does a non-finite proposal.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .qdsim import EnsembleQDsim, StepAborted

NFEAT = 6
FEAT_SCALE = 2.0 ** 32


def decode_features(acc: np.ndarray, ncell: int):
    """Finish the integer slots written by pq3d_features (nb x 4 int64):
    max log10 V, min log10 V, mean log10(V theta/Dc), fraction V > 1e-3."""
    acc = np.ascontiguousarray(acc, dtype=np.int64)
    vmax = acc[:, 0].view(np.float64) - 1000.0
    vmin = 1000.0 - acc[:, 1].view(np.float64)
    mean = acc[:, 2].astype(np.float64) / FEAT_SCALE / ncell
    frac = acc[:, 3].astype(np.float64) / ncell
    return vmax, vmin, mean, frac


class StepMLP:
    """Parameters, normalisation and Adam state on the device."""

    def __init__(self, device="cuda", seed=0):
        self.device = torch.device(device)
        self.P = int(_lib.load().pq3d_mlp_num_params())
        rng = np.random.default_rng(seed)
        p = np.zeros(self.P, dtype=np.float32)
        o = 0
        for fan_in, fan_out in ((6, 64), (64, 64), (64, 64), (64, 1)):
            k = fan_in * fan_out
            p[o:o + k] = rng.normal(0, 1.0 / np.sqrt(fan_in), k)
            o += k + fan_out  # biases stay zero
        self.prm = torch.from_numpy(p).to(self.device)
        self.m = torch.zeros(self.P, dtype=torch.float32, device=self.device)
        self.v = torch.zeros(self.P, dtype=torch.float32, device=self.device)
        self.t = 0
        self.x_mu = np.zeros(NFEAT, dtype=np.float32)
        self.x_sd = np.ones(NFEAT, dtype=np.float32)
        self.y_mu, self.y_sd = 0.0, 1.0

    # ------------------------------------------------------------ kernels
    def _fb(self, X, T=None, chunk=64, out=None):
        n = X.shape[0]
        nblk = (n + chunk - 1) // chunk
        grad = loss = None
        if T is not None:
            grad = torch.empty(nblk * self.P, dtype=torch.float32, device=self.device)
            loss = torch.empty(nblk, dtype=torch.float32, device=self.device)
        _lib.call("pq3d_mlp_forward_backward", _lib.ptr(self.prm), _lib.ptr(X), _lib.ptr(T), n,
                  chunk, _lib.ptr(out), _lib.ptr(grad), _lib.ptr(loss), _lib.stream_handle())
        return grad, loss, nblk

    def loss_and_grad(self, X, T, chunk=64):
        """Mean loss and gradient on device-normalised inputs (no update)."""
        grad, loss, nblk = self._fb(X, T, chunk)
        g = torch.empty(self.P, dtype=torch.float32, device=self.device)
        _lib.call("pq3d_mlp_adam", _lib.ptr(self.prm), _lib.ptr(grad), nblk, X.shape[0],
                  _lib.ptr(self.m), _lib.ptr(self.v), 0.0, 0.9, 0.999, 1e-8, 1, _lib.ptr(g),
                  _lib.stream_handle())
        return float(loss.sum()) / X.shape[0], g

    def predict_raw(self, X):
        out = torch.empty(X.shape[0], dtype=torch.float32, device=self.device)
        self._fb(X, None, chunk=64, out=out)
        return out

    # ---------------------------------------------------------- training
    def fit(self, feats: np.ndarray, target: np.ndarray, epochs=20, batch=4096, lr=2e-3,
            seed=0, log=None):
        feats = np.asarray(feats, dtype=np.float32)
        target = np.asarray(target, dtype=np.float32)
        self.x_mu = feats.mean(0)
        self.x_sd = feats.std(0) + 1e-6
        self.y_mu = float(target.mean())
        self.y_sd = float(target.std() + 1e-6)
        Xn = torch.from_numpy((feats - self.x_mu) / self.x_sd).to(self.device)
        Tn = torch.from_numpy((target - self.y_mu) / self.y_sd).to(self.device)
        n = Xn.shape[0]
        rng = np.random.default_rng(seed)
        hist = []
        for ep in range(epochs):
            order = torch.from_numpy(rng.permutation(n)).to(self.device)
            tot = 0.0
            for s in range(0, n, batch):
                idx = order[s:s + batch]
                xb = Xn[idx].contiguous()
                tb = Tn[idx].contiguous()
                grad, loss, nblk = self._fb(xb, tb)
                self.t += 1
                _lib.call("pq3d_mlp_adam", _lib.ptr(self.prm), _lib.ptr(grad), nblk,
                          xb.shape[0], _lib.ptr(self.m), _lib.ptr(self.v), float(lr), 0.9,
                          0.999, 1e-8, self.t, None, _lib.stream_handle())
                tot += float(loss.sum())
            hist.append(tot / n)
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.4e}")
        return hist

    def predict_log_dt(self, feats: np.ndarray) -> np.ndarray:
        X = torch.from_numpy(((np.asarray(feats, np.float32) - self.x_mu) / self.x_sd)
                             .astype(np.float32)).to(self.device)
        y = self.predict_raw(X).cpu().numpy().astype(np.float64)
        return y * self.y_sd + self.y_mu

    def state_dict(self):
        return dict(prm=self.prm.cpu().numpy(), x_mu=self.x_mu, x_sd=self.x_sd,
                    y=np.array([self.y_mu, self.y_sd]))

    def load_state_dict(self, d):
        self.prm = torch.from_numpy(np.asarray(d["prm"], np.float32)).to(self.device)
        self.x_mu, self.x_sd = np.asarray(d["x_mu"]), np.asarray(d["x_sd"])
        self.y_mu, self.y_sd = float(d["y"][0]), float(d["y"][1])


class FeatureEnsemble(EnsembleQDsim):
    """EnsembleQDsim that records per-instance features and, optionally,
    proposes step sizes with a StepMLP (learned dt)."""

    def __init__(self, problems, op, mlp: StepMLP | None = None, growth_cap=8.0, **kw):
        super().__init__(problems, op, **kw)
        self.mlp = mlp
        self.growth_cap = growth_cap
        self.acc = torch.zeros(4 * self.nb, dtype=torch.int64, device=self.device)
        self.V0 = float(problems[0].V0)
        self.last_dt = np.full(self.nb, 1e-3)
        self.last_ratio = np.full(self.nb, 1e-2)
        self.prev_err = np.full(self.nb, 1.0)
        self.log_feats, self.log_target = [], []
        self.record = False
        self.nproposals_rejected = 0

    def features(self) -> np.ndarray:
        _lib.call("pq3d_features", self.ne, self.nb, _lib.ptr(self.slipv), _lib.ptr(self.state),
                  _lib.ptr(self.b), _lib.ptr(self.f0), self.V0, _lib.ptr(self.acc),
                  _lib.stream_handle())
        a = self.acc.cpu().numpy().reshape(self.nb, 4)
        vmax, vmin, mean, frac = decode_features(a, self.n)
        return np.stack([vmax, vmin, mean, np.log10(self.last_dt), np.log10(self.last_ratio),
                         frac], axis=1)

    def step(self):
        """One accepted step per instance; first attempt uses the proposal."""
        f = self.features() if (self.mlp is not None or self.record) else None
        self.compute_K()
        h = self.dt.copy()
        if self.mlp is not None:
            prop = 10.0 ** self.mlp.predict_log_dt(f)
            ok = np.isfinite(prop) & (prop > 0)
            h = np.where(ok, np.minimum(prop, self.growth_cap * self.last_dt), h)
        tol = self.reltol
        done = np.zeros(self.nb, dtype=bool)
        nrj = np.zeros(self.nb, dtype=np.int64)
        dtnext = np.zeros(self.nb)
        target = np.zeros(self.nb)
        ratio = np.zeros(self.nb)
        while not done.all():
            e = self._attempt(h)
            for b in np.where(~done)[0]:
                R1 = np.float64(e[b, 0]) + 1e-10
                R2 = max(np.float64(e[b, 1]), np.float64(e[b, 2]), np.float64(e[b, 3])) + 1e-10
                c1, c2 = R1 / tol, R2 / tol
                hn1 = h[b] * 0.9 * (tol / R1) ** 0.2
                hn2 = h[b] * 0.9 * (tol / R2) ** 0.2
                if max(c1, c2) < 1.0 and not (np.isnan(c1) or np.isnan(c2)):
                    dtnext[b] = min(min(hn1, hn2), 1.5 * h[b])
                    target[b] = min(hn1, hn2)
                    ratio[b] = max(c1, c2)
                    done[b] = True
                else:
                    nrj[b] += 1
                    if self.mlp is not None and nrj[b] == 1:
                        self.nproposals_rejected += 1
                    err = max(c1, c2)
                    if self.mlp is not None and np.isfinite(err):
                        # PI fallback (Gustafsson): h (1/err)^0.14 (prev/1)^0.08, 0.9 safety
                        hpi = h[b] * 0.9 * (1.0 / err) ** 0.14 * (self.prev_err[b]) ** 0.08
                        h[b] = min(max(0.2 * h[b], hpi), 0.9 * h[b])
                    else:
                        h[b] = max(0.5 * h[b], min(hn1, hn2))
                    if h[b] < 1e-15 or nrj[b] > 20:
                        raise StepAborted(f"instance {b}: dt is too small (h={h[b]:g})")
        self.nrejects += nrj
        if self.record:
            self.log_feats.append(f)
            self.log_target.append(np.log10(target))
        self.hdev.copy_(torch.from_numpy(h))
        _lib.call("pq3d_rk45_commit", self.ne, _lib.ptr(self.inst), _lib.ptr(self.hdev), 0,
                  _lib.ptr(self.a), _lib.ptr(self.cellpar), _lib.ptr(self.y[0]),
                  _lib.ptr(self.y[1]), _lib.ptr(self.y[2]), _lib.ptr(self.y[3]),
                  _lib.ptr(self.Tno), _lib.ptr(self.Tt1o), _lib.ptr(self.Tt2o),
                  _lib.ptr(self.state), _lib.ptr(self.slipv1), _lib.ptr(self.slipv2),
                  _lib.ptr(self.slipv), _lib.ptr(self.rake), _lib.ptr(self.slip1),
                  _lib.ptr(self.slip2), _lib.ptr(self.vmax), self.nb, _lib.stream_handle())
        self.time += h
        self.dt = dtnext
        self.last_dt = h.copy()
        self.last_ratio = np.maximum(ratio, 1e-12)
        self.prev_err = np.maximum(ratio, 1e-6)
        self.nsteps += 1
        return h

    def dataset(self):
        return np.concatenate(self.log_feats), np.concatenate(self.log_target)
