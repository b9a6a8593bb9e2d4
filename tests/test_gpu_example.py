"""The reference's example models, from their committed archives
(tests/golden/make_model_fixtures.py), run on the device: BP5-QD (9,250
triangles, half space, 3e-2 m/s nucleation disc), Lab-model and
EAFZ-model."""
import os

import numpy as np
import pytest
import torch

from oracle.qdsim_oracle import OracleSim
from pyquake3d_amd.hmatrix import HMatrix
from pyquake3d_amd.problem import load_problem
from pyquake3d_amd.qdsim import DenseOperator, DeviceEnsemble, QDsim

pytestmark = pytest.mark.gpu
ARCHIVE = os.path.join(os.path.dirname(__file__), "golden", "bp5_example.npz")
YEAR = 3.15576e7


def test_bp5_example_steps_match_oracle(gpu, pq3d_lib):
    """Dense FP64 half-space operators built on the device; 25 steps of the
    device stepper against the CPU oracle of the reference step on the same
    operators (slip rates up to ~1 m/s in the nucleation disc): step sizes
    to 1e-6, slip rates to 1e-6 of the largest (still above 1e-2 m/s)."""
    p, hs = load_problem(ARCHIVE)
    assert hs and p.n == 9250
    op = DenseOperator(p.mesh, p.mu, p.lambda_, device=gpu, halfspace=hs)
    M = op.matrix().cpu().numpy()
    n = p.n

    def mv(v1, v2):
        y = M @ np.concatenate([v1, v2])
        return y[:n], y[n:2 * n], y[2 * n:]

    sim = QDsim(p, op, device=gpu)
    ora = OracleSim(None, p.a, p.b, p.dc, p.f0, p.V0, p.mu, p.Cs, p.Vpl, p.rake,
                    p.initial_state(), shear_loading=p.shear_loading,
                    normal_loading=p.normal_loading, fix_Tn=p.fix_Tn, matvec=mv)
    dg = do = 1e-3
    for i in range(25):
        hg, dg = sim.simu_forward(dg)
        ho, do = ora.simu_forward(do)
        assert hg == pytest.approx(ho, rel=1e-6), i
    vg = sim.snapshot()["slipv"]
    assert np.abs(vg - ora.slipv).max() <= 1e-6 * np.abs(ora.slipv).max()
    assert ora.slipv.max() > 1e-2         # still inside the seismic phase


def test_bp5_example_fp32_hmatrix_tracks_dense_fp64(gpu, pq3d_lib):
    """FP32 half-space H-matrix (leaf 32, eta 2, ACA eps 1e-6) through the
    first rupture: after 4,000 accepted steps both runs have passed a peak
    slip rate above 0.5 m/s without tripping the reference's abort rule, and
    their simulated times agree to 1 %."""
    p, hs = load_problem(ARCHIVE)
    runs = []
    for op in (DenseOperator(p.mesh, p.mu, p.lambda_, device=gpu, halfspace=hs),
               HMatrix.from_mesh(p.mesh, p.mu, p.lambda_, leaf=32, eta=2.0, eps=1e-6,
                                 dtype=torch.float32, device=gpu, halfspace=hs)):
        ens = DeviceEnsemble([p], op, device=gpu, rounds=8)
        peak = 0.0
        while True:
            ens.step()
            ens.sync()
            peak = max(peak, float(ens.max_slipv()[0]))
            if ens.naccepted[0] >= 4000:
                break
        runs.append((ens.time[0], peak))
        del op, ens
        torch.cuda.empty_cache()
    (t64, peak64), (t32, peak32) = runs
    assert peak64 > 0.5 and peak32 > 0.5
    assert abs(t32 - t64) <= 0.01 * t64


def _run_model(p, hs, dtype, gpu, steps, vth=None):
    """(simulated time after ``steps`` accepted steps, time at which max V
    first exceeded vth or None, peak max V) on an H-matrix (leaf 32, eta 2,
    ACA eps 1e-6) with the device controller."""
    H = HMatrix.from_mesh(p.mesh, p.mu, p.lambda_, leaf=32, eta=2.0, eps=1e-6, dtype=dtype,
                          device=gpu, halfspace=hs)
    ens = DeviceEnsemble([p], H, device=gpu, rounds=8)
    t_cross, peak = None, 0.0
    while ens.naccepted[0] < steps:
        ens.step()
        ens.sync()
        v = float(ens.max_slipv()[0])
        assert np.isfinite(v)
        peak = max(peak, v)
        if vth is not None and t_cross is None and v > vth:
            t_cross = float(ens.time[0])
    return float(ens.time[0]), t_cross, peak


def test_lab_example_runs_fp32_and_fp64(gpu, pq3d_lib):
    """The reference's laboratory model (examples/Lab-model: 14,336
    triangles, 4 cm dipping fault, full space, Initlab normal stress, Dc
    1e-7 m): 300 accepted steps from its initial state in FP64 and FP32
    storage, no abort, simulated times within 10 % (the step sizes of the
    stiff lab-scale problem follow the operator's rounding)."""
    p, hs = load_problem(os.path.join(os.path.dirname(__file__), "golden", "lab_example.npz"))
    assert not hs and p.n == 14336
    assert 45.0 < p.Tno.min() < p.Tno.max() < 75.0      # edge taper and smoothed noise
    t64, _, peak64 = _run_model(p, hs, torch.float64, gpu, 300)
    t32, _, peak32 = _run_model(p, hs, torch.float32, gpu, 300)
    assert t64 > 0 and abs(t32 - t64) <= 0.1 * t64
    assert 3e-5 <= peak64 < 1e-2 and 3e-5 <= peak32 < 1e-2


def test_eafz_example_nucleation_fp32_tracks_fp64(gpu, pq3d_lib):
    """The reference's East Anatolian fault model (examples/EAFZ-model:
    41,800 triangles, half space, friction zones from the principal stresses
    and the mesh boundary): 300 accepted steps from the initial state in
    FP64 and FP32 storage; both nucleate (max V through 1e-6 m/s) and the
    times at which they do agree to 1 %."""
    p, hs = load_problem(os.path.join(os.path.dirname(__file__), "golden", "eafz_example.npz"))
    assert hs and p.n == 41800
    t64, c64, peak64 = _run_model(p, hs, torch.float64, gpu, 300, vth=1e-6)
    t32, c32, peak32 = _run_model(p, hs, torch.float32, gpu, 300, vth=1e-6)
    assert c64 is not None and c32 is not None
    assert abs(c32 - c64) <= 0.01 * c64
    assert peak64 > 1e-5 and peak32 > 1e-5
