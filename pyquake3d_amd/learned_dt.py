"""Learned step-size proposal for the RK45 stepper.

A 6 -> 64 -> 64 -> 64 -> 1 MLP (FP32, tanh) corrects the step the reference
controller would try next. Inputs per instance:
  [max log10 V, min log10 V, mean log10(V theta/Dc), fraction of cells with
   V > 1e-3 m/s, log10 of the accepted attempt's error ratio max(c1, c2),
   log10(dtnext / h) of the reference controller]
(the cell reductions are one HIP kernel, ``pq3d_features``). The target is
log10(h_accepted / dtnext): how far the reference controller
(src/QDsim.py:931-950) ended up shrinking its own proposal dtnext through
rejections. The network is trained on the GPU with the library's forward,
backward and Adam kernels on (features, target) pairs logged from runs of the
reference controller.

In ``FeatureEnsemble`` with a network every step still goes through the
reference's embedded error test: the first attempt uses
dtnext * 10^min(prediction - margin, 0); a rejected attempt falls back to a
PI controller (Gustafsson: h (1/err)^0.14 (err_prev)^0.08 with 0.9 safety,
clamped to [0.2 h, 0.9 h]), as does a non-finite prediction. The reference's
dtnext rule (0.9 h (tol/err)^0.2, at most 1.5 h) is unchanged.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .qdsim import EnsembleQDsim, StepAborted

NFEAT = 6
FEAT_SCALE = 2.0 ** 32


def decode_features(acc: np.ndarray, ncell):
    """Finish the integer slots written by pq3d_features (nb x 4 int64):
    max log10 V, min log10 V, mean log10(V theta/Dc), fraction V > 1e-3.
    ``ncell``: cells per instance (a number, or one per instance)."""
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
        # normalisation in FP32 values: the device proposal kernel reads them as FP32
        self.x_mu = feats.mean(0).astype(np.float32)
        self.x_sd = (feats.std(0) + 1e-6).astype(np.float32)
        self.y_mu = float(np.float32(target.mean()))
        self.y_sd = float(np.float32(target.std() + 1e-6))
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
        self.x_mu = np.asarray(d["x_mu"], np.float32)
        self.x_sd = np.asarray(d["x_sd"], np.float32)
        self.y_mu = float(np.float32(d["y"][0]))
        self.y_sd = float(np.float32(d["y"][1]))

    def save(self, path, **meta) -> None:
        """Parameters and normalisation as one .npz (no pickles); ``meta``
        strings are stored alongside (e.g. how the network was trained)."""
        d = self.state_dict()
        np.savez(path, **d, **{"meta_" + k: np.array(str(v)) for k, v in meta.items()})

    @classmethod
    def load(cls, path, device="cuda"):
        with np.load(path, allow_pickle=False) as z:
            d = {k: np.array(z[k]) for k in z.files}
        if d["prm"].shape != (int(_lib.load().pq3d_mlp_num_params()),):
            raise ValueError(f"{path}: parameter count does not match the network")
        m = cls(device=device)
        m.load_state_dict(d)
        m.meta = {k[5:]: str(v) for k, v in d.items() if k.startswith("meta_")}
        return m


class FeatureEnsemble(EnsembleQDsim):
    """EnsembleQDsim that records per-instance features and targets and,
    with a trained StepMLP, proposes the first attempt's step size."""

    def __init__(self, problems, op, mlp: StepMLP | None = None, margin=0.0, **kw):
        super().__init__(problems, op, **kw)
        self.mlp = mlp
        self.margin = margin
        self.acc = torch.zeros(4 * self.nb, dtype=torch.int64, device=self.device)
        self.V0 = float(problems[0].V0)
        self.last_ratio = np.full(self.nb, 0.5)      # max(c1, c2) at the last acceptance
        self.last_growth = np.ones(self.nb)          # dtnext / h at the last acceptance
        self.log_feats, self.log_target = [], []
        self.record = False
        self.nfirst_rejected = 0                      # steps whose first attempt was rejected
        self.nfallback = 0                            # PI fallback attempts
        self.max_accepted_ratio = 0.0                 # largest max(c1, c2) ever accepted

    def features(self) -> np.ndarray:
        _lib.call("pq3d_features", self.ne, self.nb, _lib.ptr(self.inst), _lib.ptr(self.slipv),
                  _lib.ptr(self.state), _lib.ptr(self.b), _lib.ptr(self.f0), self.V0,
                  _lib.ptr(self.acc), _lib.stream_handle())
        a = self.acc.cpu().numpy().reshape(self.nb, 4)
        vmax, vmin, mean, frac = decode_features(a, np.diff(self.off))
        return np.stack([vmax, vmin, mean, frac, np.log10(self.last_ratio),
                         np.log10(self.last_growth)], axis=1)

    def step(self):
        """One accepted step per instance."""
        f = self.features() if (self.mlp is not None or self.record) else None
        self.compute_K()
        base = self.dt.copy()                         # the reference controller's dtnext
        h = base.copy()
        learned = np.zeros(self.nb, dtype=bool)
        if self.mlp is not None:
            pred = self.mlp.predict_log_dt(f)
            learned = np.isfinite(pred)
            corr = np.where(learned, np.minimum(pred - self.margin, 0.0), 0.0)
            h = np.where(learned, base * 10.0 ** corr, base)
        tol = self.reltol
        done = np.zeros(self.nb, dtype=bool)
        nrj = np.zeros(self.nb, dtype=np.int64)
        dtnext = np.zeros(self.nb)
        ratio = np.ones(self.nb)
        prev = np.maximum(self.last_ratio, 1e-6)
        while not done.all():
            e = self._attempt(np.where(done, -1.0, h))
            for b in np.where(~done)[0]:
                R1 = np.float64(e[b, 0]) + 1e-10
                R2 = max(np.float64(e[b, 1]), np.float64(e[b, 2]), np.float64(e[b, 3])) + 1e-10
                c1, c2 = R1 / tol, R2 / tol
                hn1 = h[b] * 0.9 * (tol / R1) ** 0.2
                hn2 = h[b] * 0.9 * (tol / R2) ** 0.2
                if max(c1, c2) < 1.0 and not (np.isnan(c1) or np.isnan(c2)):
                    dtnext[b] = min(min(hn1, hn2), 1.5 * h[b])
                    ratio[b] = max(c1, c2)
                    self.max_accepted_ratio = max(self.max_accepted_ratio, ratio[b])
                    done[b] = True
                    continue
                nrj[b] += 1
                if nrj[b] == 1:
                    self.nfirst_rejected += 1
                err = max(c1, c2)
                if learned[b] and np.isfinite(err):
                    # PI fallback (Gustafsson), 0.9 safety, within [0.2 h, 0.9 h]
                    self.nfallback += 1
                    hpi = h[b] * 0.9 * (1.0 / err) ** 0.14 * prev[b] ** 0.08
                    h[b] = min(max(0.2 * h[b], hpi), 0.9 * h[b])
                else:
                    h[b] = max(0.5 * h[b], min(hn1, hn2))
                if h[b] < 1e-15 or nrj[b] > 20:
                    raise StepAborted(f"instance {b}: dt is too small (h={h[b]:g})")
        self.nrejects += nrj
        if self.record:
            self.log_feats.append(f)
            self.log_target.append(np.log10(h / base))
        self.hdev.copy_(torch.from_numpy(h))
        _lib.call("pq3d_rk45_commit", self.ne, _lib.ptr(self.inst), _lib.ptr(self.hdev), 0,
                  _lib.ptr(self.a), _lib.ptr(self.cellpar), _lib.ptr(self.y[0]),
                  _lib.ptr(self.y[1]), _lib.ptr(self.y[2]), _lib.ptr(self.y[3]),
                  _lib.ptr(self.Tno), _lib.ptr(self.Tt1o), _lib.ptr(self.Tt2o),
                  _lib.ptr(self.state), _lib.ptr(self.slipv1), _lib.ptr(self.slipv2),
                  _lib.ptr(self.slipv), _lib.ptr(self.rake), _lib.ptr(self.slip1),
                  _lib.ptr(self.slip2), _lib.ptr(self.vmax), self.nb, _lib.stream_handle())
        self.time += h
        self.last_growth = dtnext / h
        self.dt = dtnext
        self.last_ratio = np.maximum(ratio, 1e-12)
        self.nsteps += 1
        return h

    def dataset(self):
        return np.concatenate(self.log_feats), np.concatenate(self.log_target)

    def save_state(self) -> dict:
        d = super().save_state()
        d["last_ratio"], d["last_growth"] = self.last_ratio.copy(), self.last_growth.copy()
        return d

    def load_state(self, d: dict) -> None:
        super().load_state(d)
        self.last_ratio, self.last_growth = d["last_ratio"].copy(), d["last_growth"].copy()
