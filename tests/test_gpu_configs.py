"""The time-stepping configurations at their stated sizes.

C3: H-matrix RK45 stepping at N = 100,352 (the headline's draws: BP5-like
    friction, Dc x 0.7, normal traction fixed), FP32 against FP64 operator
    storage, single runs and the 16-draw matrix-core ensemble.
C4: 64 simulations on meshes of uneven size (N log-uniform in
    [2,000, 20,000]) as one block-diagonal FP32 H-matrix: every 8th equals its
    FP32 solo run bit for bit and its first event onset agrees with its FP64
    solo run within 1 %; the smallest, on dense FP64 operators, follows the
    CPU oracle of the reference step (src/QDsim.py:892-1134).
"""
from dataclasses import replace

import numpy as np
import pytest
import torch

from pyquake3d_amd.hmatrix import HMatrix, MultiHMatrix
from pyquake3d_amd.problem import ensemble_c4, undulating_c3
from pyquake3d_amd.qdsim import DeviceEnsemble

pytestmark = pytest.mark.gpu


def _c3_draws(nb):
    p0 = undulating_c3()
    draws = [p0] + [undulating_c3(friction_seed=k, mesh=p0.mesh) for k in range(1, nb)]
    return [replace(p, dc=p.dc * 0.7, fix_Tn=True) for p in draws]


def _advance(ens, nacc):
    """Iterate until every instance has ``nacc`` accepted steps (an instance
    that gets there first is marked finished, flag 3, and takes no more)."""
    for _ in range(20 * nacc):
        ens.step()
        ens.sync()
        done = ens.naccepted >= nacc
        if done.any():
            fl = ens.flag.cpu()
            fl[torch.from_numpy(done)] = 3
            ens.flag.copy_(fl)
        if done.all():
            return
    raise AssertionError("instances did not reach the step count")


def test_c3_first_event_fp32_against_fp64_storage(gpu, pq3d_lib):
    """Draw 0 at N = 100,352 from its seeded initial state through its first
    seismic event (max slip rate through 1e-2 m/s at ~0.97 yr) with the
    device controller: FP32 storage with FP64 sums, and the 16-draw
    matrix-core ensemble, against FP64 storage. The event onset agrees within
    0.5 % and the peak slip rate within 5 %. (Step for step the runs are not
    comparable once the patch accelerates: a perturbation of the operator
    shifts the rupture by a fraction of a step.)"""
    draws = _c3_draws(16)
    p = draws[0]
    assert p.n == 100352
    year = 3.15576e7
    res = {}
    for name, dtype, acc, probs in (("fp64", torch.float64, "f64", [p]),
                                    ("fp32", torch.float32, "f64", [p]),
                                    ("fp32_mfma16", torch.float32, "mfma", draws)):
        H = HMatrix.from_mesh(p.mesh, p.mu, p.lambda_, leaf=64, eta=2.0, eps=1e-4, dtype=dtype,
                              device=gpu)
        H.accumulate = acc
        ens = DeviceEnsemble(probs, H, device=gpu, rounds=8, t_end=1.2 * year, vup=1e-2,
                             vdown=1e-4)
        for i in range(200000):
            ens.step()
            if i % 200 == 199 and ens.finished()[0]:
                break
        ens.sync()
        assert ens.finished()[0] and not ens.aborted()[0]
        ev = ens.events()[0]
        assert len(ev) >= 1, name
        res[name] = (float(ev[0]) / year, float(ens.peak_slipv()[0]))
        del H, ens
        torch.cuda.empty_cache()
    t64, v64 = res["fp64"]
    assert 0.5 < t64 < 1.2 and v64 > 0.1
    for name in ("fp32", "fp32_mfma16"):
        t, v = res[name]
        print(f"{name}: first onset {t:.5f} yr (FP64 {t64:.5f}), peak V {v:.4f} ({v64:.4f})")
        assert abs(t - t64) / t64 < 5e-3 and abs(v - v64) / v64 < 5e-2, (name, t, t64, v, v64)


def _to_end(ens, limit=60000):
    for i in range(limit):
        ens.step()
        if i % 200 == 199 and (ens.finished() | ens.aborted()).all():
            break
    ens.sync()
    assert ens.finished().all()


def test_c4_ragged_ensemble_at_size(gpu, pq3d_lib, oracle):
    """64 instances on one block-diagonal FP32 H-matrix (total ~0.58 M
    cells) integrated over [0, 100] yr with the device controller. Every 8th
    instance run alone in FP32 ends at the same time with the same event
    onsets bit for bit; run alone in FP64 storage it has the same number of
    events (max slip rate through 1e-2 m/s) with onsets within 1 % (two of
    the eight have events in the window: onsets within 0.03 %; the others
    creep or load without a seismic event)."""
    from oracle.qdsim_oracle import OracleSim, dense_operators
    from pyquake3d_amd.mesh import get_eleVec
    from pyquake3d_amd.qdsim import DenseOperator, QDsim
    year = 3.15576e7
    kw = dict(device=gpu, rounds=8, t_end=100 * year, vup=1e-2, vdown=1e-4)
    probs = ensemble_c4(64, seed=0)
    sizes = np.array([q.n for q in probs])
    assert len(probs) == 64 and sizes.min() >= 1500 and sizes.max() <= 21000
    hs = [HMatrix.from_mesh(q.mesh, q.mu, q.lambda_, leaf=64, eta=2.0, eps=1e-4,
                            dtype=torch.float32, device=gpu) for q in probs]
    ens = DeviceEnsemble(probs, MultiHMatrix(hs), **kw)
    _to_end(ens)
    ev = ens.events()
    compared = 0
    for k in range(0, 64, 8):
        solo = DeviceEnsemble([probs[k]], hs[k], **kw)
        _to_end(solo)
        assert solo.time[0] == ens.time[k], k
        assert np.array_equal(solo.events()[0], ev[k]), k
        H64 = HMatrix.from_mesh(probs[k].mesh, probs[k].mu, probs[k].lambda_, leaf=64, eta=2.0,
                                eps=1e-4, dtype=torch.float64, device=gpu)
        s64 = DeviceEnsemble([probs[k]], H64, **kw)
        _to_end(s64)
        e64 = s64.events()[0]
        print(f"instance {k} (N={probs[k].n}): onsets FP32 {np.round(ev[k] / year, 4)} "
              f"FP64 {np.round(e64 / year, 4)}")
        assert len(e64) == len(ev[k]), k
        if len(e64):
            assert (np.abs(ev[k] - e64) / e64 < 1e-2).all(), k
            compared += 1
        del H64, s64
    assert compared >= 2
    # the smallest instance on dense FP64 operators against the CPU oracle
    k = int(np.argmin(sizes))
    q = probs[k]
    ev, xg = get_eleVec(q.mesh.nodelst, q.mesh.elelst)
    K = dense_operators(oracle, xg, ev, q.mesh.triangles(), q.mu, q.lambda_)
    sim = QDsim(q, DenseOperator(q.mesh, q.mu, q.lambda_, device=gpu), device=gpu)
    ora = OracleSim(K, q.a, q.b, q.dc, q.f0, q.V0, q.mu, q.Cs, q.Vpl, q.rake, q.initial_state())
    dg = do = 1e-3
    for i in range(100):
        hg, dg = sim.simu_forward(dg)
        ho, do = ora.simu_forward(do)
        if i < 50:
            assert hg == pytest.approx(ho, rel=1e-7), i
    assert sim.time == pytest.approx(ora.time, rel=1e-6)
