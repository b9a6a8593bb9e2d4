"""GPU time step (fused RK45 attempt, controller, commit) against the CPU
oracle of the reference step on the same dense FP64 operators."""
import numpy as np
import pytest
import torch

from oracle.qdsim_oracle import OracleSim, dense_operators
from pyquake3d_amd.mesh import get_eleVec
from pyquake3d_amd.problem import planar_c1
from pyquake3d_amd.qdsim import DenseOperator, QDsim

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(oracle):
    p = planar_c1(nx=12, nz=6, seed=0)
    ev, xg = get_eleVec(p.mesh.nodelst, p.mesh.elelst)
    K = dense_operators(oracle, xg, ev, p.mesh.triangles(), p.mu, p.lambda_)
    return p, K


def _pair(p, K, gpu, tableau="ref"):
    op = DenseOperator(p.mesh, p.mu, p.lambda_, device=gpu)
    sim = QDsim(p, op, device=gpu, tableau=tableau)
    ora = OracleSim(K, p.a, p.b, p.dc, p.f0, p.V0, p.mu, p.Cs, p.Vpl, p.rake, p.initial_state(),
                    tableau=tableau)
    return sim, ora


def test_dense_operator_matches_oracle(gpu, pq3d_lib, small):
    p, K = small
    op = DenseOperator(p.mesh, p.mu, p.lambda_, device=gpu)
    assert op.H.stats.n_lowrank == 0
    rng = np.random.default_rng(0)
    v1, v2 = rng.standard_normal(p.n), rng.standard_normal(p.n)
    pm = op.perm
    y = op.matvec(torch.from_numpy(v1[pm]).to(gpu), torch.from_numpy(v2[pm]).to(gpu))
    yo = np.empty((3, p.n))
    yo[:, pm] = y.cpu().numpy()
    ref = np.stack([K["A1s"] @ v1 + K["A1d"] @ v2, K["A2s"] @ v1 + K["A2d"] @ v2,
                    K["Bs"] @ v1 + K["Bd"] @ v2])
    assert np.abs(yo - ref).max() <= 1e-12 * np.abs(ref).max()
    # the dense matrix of the operator, in original order
    M = op.matrix().cpu().numpy()
    assert np.abs(M @ np.concatenate([v1, v2]) - ref.ravel()).max() <= 1e-12 * np.abs(ref).max()


def test_derivative_matches_oracle(gpu, pq3d_lib, small):
    p, K = small
    sim, ora = _pair(p, K, gpu)
    sim.compute_K()
    ora.simu_forward(1e-3)  # sets ora.AdotV1 ... (state before the step is what we compare)
    ora2 = OracleSim(K, p.a, p.b, p.dc, p.f0, p.V0, p.mu, p.Cs, p.Vpl, p.rake,
                     p.initial_state())
    ora2.AdotV1, ora2.AdotV2, ora2.dsigmadt = ora.AdotV1, ora.AdotV2, ora.dsigmadt
    rng = np.random.default_rng(1)
    Tno = ora2.Tno * (1 + 0.01 * rng.standard_normal(p.n))
    Tt1 = ora2.Tt1o * (1 + 0.01 * rng.standard_normal(p.n))
    Tt2 = 0.01 * ora2.Tt1o * rng.standard_normal(p.n)
    st = ora2.state + 0.001 * rng.standard_normal(p.n)
    ref = ora2.derivative(Tno, Tt1, Tt2, st)
    pm = sim.perm      # the stepper works in the operator's cluster order
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x[pm])).to(gpu)
    got = [x.cpu().numpy() for x in sim.derivative(d(Tno), d(Tt1), d(Tt2), d(st))]
    for g, r in zip(got, ref):
        assert np.abs(g - r[pm]).max() <= 1e-12 * np.abs(r).max() + 1e-300


@pytest.mark.parametrize("tableau", ["ref", "dp5"])
def test_trajectory_matches_oracle(gpu, pq3d_lib, small, tableau):
    """Accepted steps from the seeded nucleation state. Rounding differences
    (device exp, GEMV order) are amplified by the controller's err^-0.2, so the
    comparison is tight over the first 80 steps and looser after. With the
    published tableau the step sizes grow until the frozen-operator coupling
    makes the run sensitive to rounding, so only the first 80 steps are
    compared for it."""
    p, K = small
    sim, ora = _pair(p, K, gpu, tableau)
    dt_g = dt_o = 1e-3
    for i in range(80):
        hg, dt_g = sim.simu_forward(dt_g)
        ho, dt_o = ora.simu_forward(dt_o)
        assert hg == pytest.approx(ho, rel=1e-7), i
    snap = sim.snapshot()
    for k in ("Tno", "Tt1o", "state", "slipv"):
        ref = getattr(ora, k)
        assert np.abs(snap[k] - ref).max() <= 1e-7 * np.abs(ref).max(), k
    if tableau == "dp5":
        return
    for i in range(220):
        _, dt_g = sim.simu_forward(dt_g)
        _, dt_o = ora.simu_forward(dt_o)
    assert sim.time == pytest.approx(ora.time, rel=1e-3)


def _events(vmax, t, thr=1e-3):
    ev, inev = [], False
    for v, tt in zip(vmax, t):
        if v > thr and not inev:
            inev = True
            ev.append([tt, v])
        if inev:
            ev[-1][1] = max(ev[-1][1], v)
            if v < thr / 10:
                inev = False
    return ev


def test_event_times_and_peak_slip_rate_match_oracle(gpu, pq3d_lib, small):
    """Accuracy criterion on the 144-cell problem: the first three seismic
    events (max V > 1e-3 m/s) agree with the CPU oracle of the reference step
    within 1% in onset time. The peak slip rate is the largest value at the
    accepted steps, so it depends on where the steps fall within the rupture;
    it must agree within 0.1 in log10 V (26%)."""
    p, K = small
    sim, ora = _pair(p, K, gpu)
    vg, tg, vo, to = [], [], [], []
    dt_g = dt_o = 1e-3
    for _ in range(4600):
        _, dt_g = sim.simu_forward(dt_g)
        _, dt_o = ora.simu_forward(dt_o)
        vg.append(sim.max_slipv())
        tg.append(sim.time)
        vo.append(float(ora.slipv.max()))
        to.append(ora.time)
    eg, eo = _events(vg, tg), _events(vo, to)
    assert len(eo) >= 3 and len(eg) >= 3, (len(eo), len(eg))
    for (t1, v1), (t2, v2) in zip(eg[:3], eo[:3]):
        assert abs(t1 - t2) <= 0.01 * t2
        assert abs(np.log10(v1) - np.log10(v2)) <= 0.1


def test_ensemble_reproduces_single_runs_bitwise(gpu, pq3d_lib):
    """Four friction draws stepped as one ensemble (one 4-RHS H-matvec per
    step) give exactly the trajectories of four separate runs: every
    right-hand side is accumulated in the same order as a single one."""
    import torch
    from dataclasses import replace

    from pyquake3d_amd.hmatrix import HMatrix
    from pyquake3d_amd.qdsim import EnsembleQDsim

    base = planar_c1(nx=16, nz=8, seed=0)
    rng = np.random.default_rng(5)
    probs = [replace(base, a=base.a * (1 + 0.05 * rng.standard_normal(base.n)),
                     dc=base.dc * rng.uniform(0.8, 1.2)) for _ in range(4)]
    H = HMatrix.from_mesh(base.mesh, base.mu, base.lambda_, leaf=16, eta=2.0, eps=1e-6,
                          dtype=torch.float32, device=gpu)
    ens = EnsembleQDsim(probs, H, device=gpu)
    singles = [QDsim(p, H, device=gpu) for p in probs]
    dts = [1e-3] * 4
    for _ in range(150):
        ens.step()
        for k, s in enumerate(singles):
            _, dts[k] = s.simu_forward(dts[k])
    for k, s in enumerate(singles):
        assert ens.time[k] == s.time
        assert np.array_equal(ens.instance("slipv", k), s.to_original(s.slipv))
        assert np.array_equal(ens.instance("state", k), s.to_original(s.state))


def test_device_controller_matches_host_controller(gpu, pq3d_lib, small):
    """The accept/reject loop run on the device (no host synchronisation)
    takes the steps of the reference loop run on the host over the same
    attempt kernel. The only difference is the rounding of pow (device
    library vs host libm), which the controller's err^-0.2 amplifies like any
    rounding difference: step sizes agree to 1e-7 over the first 80 steps,
    simulated time to 1e-5 after 300."""
    from pyquake3d_amd.qdsim import DeviceEnsemble
    p, _ = small
    op = DenseOperator(p.mesh, p.mu, p.lambda_, device=gpu)
    sim = QDsim(p, op, device=gpu)
    ens = DeviceEnsemble([p], op, device=gpu, rounds=8)
    dt = 1e-3
    t_prev = 0.0
    for i in range(300):
        h, dt = sim.simu_forward(dt)
        while True:
            ens.step()
            ens.sync()
            if ens.naccepted[0] == i + 1:
                break
        hd = ens.time[0] - t_prev
        t_prev = ens.time[0]
        if i < 80:
            assert hd == pytest.approx(h, rel=1e-7), i
    assert ens.time[0] == pytest.approx(sim.time, rel=1e-5)


def test_device_ensemble_is_batch_and_round_independent(gpu, pq3d_lib):
    """Bitwise: four friction draws batched (one 4-RHS H-matvec per step)
    equal four single runs, and carrying an unfinished step over to the next
    iteration (rounds=1) equals finishing it within one (rounds=4)."""
    from dataclasses import replace

    from pyquake3d_amd.hmatrix import HMatrix
    from pyquake3d_amd.qdsim import DeviceEnsemble

    base = planar_c1(nx=16, nz=8, seed=0)
    rng = np.random.default_rng(6)
    probs = [replace(base, a=base.a * (1 + 0.05 * rng.standard_normal(base.n)),
                     dc=base.dc * rng.uniform(0.8, 1.2)) for _ in range(4)]
    H = HMatrix.from_mesh(base.mesh, base.mu, base.lambda_, leaf=16, eta=2.0, eps=1e-6,
                          dtype=torch.float32, device=gpu)
    nacc = 200
    batch = DeviceEnsemble(probs, H, device=gpu, rounds=3)
    # step the batch until every draw has exactly nacc accepted steps: stop a
    # draw by comparing per-draw snapshots taken when it reaches nacc
    snaps = {}
    while len(snaps) < 4:
        batch.step()
        batch.sync()
        for k in range(4):
            if k not in snaps and batch.naccepted[k] == nacc:
                snaps[k] = (batch.time[k], batch.instance("slipv", k), batch.instance("state", k))
    for k, p in enumerate(probs):
        for rounds in (1, 4):
            one = DeviceEnsemble([p], H, device=gpu, rounds=rounds)
            while True:
                one.step()
                one.sync()
                if one.naccepted[0] == nacc:
                    break
            t, v, st = snaps[k]
            assert one.time[0] == t, (k, rounds)
            assert np.array_equal(one.instance("slipv", 0), v), (k, rounds)
            assert np.array_equal(one.instance("state", 0), st), (k, rounds)


def test_config1_dense_fp64_rtol_1e8(gpu, pq3d_lib, oracle):
    """The plain method at its reference size: planar 40 x 20 km fault, 1,980
    triangles, dense FP64 operators, rtol 1e-8, 1,000 RK45 steps, against the
    CPU oracle of the reference step. Step sizes agree to 1e-7 over the first
    100 steps (rounding only: device exp, GEMV order), the simulated time
    after 1,000 steps to 1e-4."""
    p = planar_c1()
    assert p.n == 1980
    ev, xg = get_eleVec(p.mesh.nodelst, p.mesh.elelst)
    K = dense_operators(oracle, xg, ev, p.mesh.triangles(), p.mu, p.lambda_)
    op = DenseOperator(p.mesh, p.mu, p.lambda_, device=gpu)
    sim = QDsim(p, op, device=gpu, reltol=1e-8)
    ora = OracleSim(K, p.a, p.b, p.dc, p.f0, p.V0, p.mu, p.Cs, p.Vpl, p.rake, p.initial_state(),
                    reltol=1e-8)
    dt_g = dt_o = 1e-3
    for i in range(1000):
        hg, dt_g = sim.simu_forward(dt_g)
        ho, dt_o = ora.simu_forward(dt_o)
        if i < 100:
            assert hg == pytest.approx(ho, rel=1e-7), i
    assert sim.time == pytest.approx(ora.time, rel=1e-4)


def test_ragged_ensemble_matches_solo_runs(gpu, pq3d_lib):
    """Eight simulations on meshes of uneven size (N from 100 to 600, planar
    or undulating, own friction draws) stepped as one block-diagonal
    operator (every matvec phase one launch for all) with the device
    controller. Each instance equals its own FP32 solo run bit for bit, and
    stays within 1e-3 in simulated time of its FP64 solo run after 150
    accepted steps."""
    from pyquake3d_amd.hmatrix import HMatrix, MultiHMatrix
    from pyquake3d_amd.problem import ensemble_c4
    from pyquake3d_amd.qdsim import DeviceEnsemble

    probs = ensemble_c4(8, seed=1, n_min=100, n_max=600)
    sizes = [p.n for p in probs]
    assert len(set(sizes)) > 3
    H32 = [HMatrix.from_mesh(p.mesh, p.mu, p.lambda_, leaf=16, eta=2.0, eps=1e-6,
                             dtype=torch.float32, device=gpu) for p in probs]
    M = MultiHMatrix(H32)
    nacc = 150
    ens = DeviceEnsemble(probs, M, device=gpu, rounds=4)
    snaps = {}
    while len(snaps) < len(probs):
        ens.step()
        ens.sync()
        for k in range(len(probs)):
            if k not in snaps and ens.naccepted[k] == nacc:
                snaps[k] = (ens.time[k], ens.instance("slipv", k))

    def solo(p, H):
        one = DeviceEnsemble([p], H, device=gpu, rounds=4)
        while True:
            one.step()
            one.sync()
            if one.naccepted[0] == nacc:
                return one.time[0], one.instance("slipv", 0)

    for k, (p, H) in enumerate(zip(probs, H32)):
        t, v = solo(p, H)
        assert t == snaps[k][0] and np.array_equal(v, snaps[k][1]), k
        H64 = HMatrix.from_mesh(p.mesh, p.mu, p.lambda_, leaf=16, eta=2.0, eps=1e-6,
                                dtype=torch.float64, device=gpu)
        t64, _ = solo(p, H64)
        assert abs(snaps[k][0] - t64) <= 1e-3 * t64, (k, snaps[k][0], t64)


def test_restore_and_vtk_restart(gpu, pq3d_lib, small, tmp_path):
    """Restore from a snapshot continues bit for bit; restart from the
    reference's VTK file (6-decimal tractions and state) continues within
    the rounding of that format; the state log gets one line per step."""
    from pyquake3d_amd import writers
    p, _ = small
    op = DenseOperator(p.mesh, p.mu, p.lambda_, device=gpu)
    a = QDsim(p, op, device=gpu)
    log = writers.StateLog(tmp_path / "state.txt", p)
    a.run(40, log=log)
    log.close(0.0)
    lines = (tmp_path / "state.txt").read_text().split("\n")
    steps = [ln for ln in lines[15:] if ln and not ln.startswith(("Program", "Time"))]
    assert len(steps) == 40 and steps[0].startswith("0 ") and steps[-1].startswith("39 ")
    snap = a.snapshot()
    a.ouputVTK(tmp_path / "step40.vtk")
    b = QDsim(p, op, device=gpu)
    b.restore(snap)
    c = QDsim(p, op, device=gpu)
    c.restore(writers.restart_fields(writers.read_vtk(tmp_path / "step40.vtk")["cells"]))
    c.time = a.time
    dt = a.htry
    for _ in range(10):
        ha, _ = a.simu_forward(dt)
        hb, _ = b.simu_forward(dt)
        hc, _ = c.simu_forward(dt)
        assert ha == hb
    sa, sb, sc = a.snapshot(), b.snapshot(), c.snapshot()
    for k in ("slipv", "Tt1o", "state", "slip1"):
        assert np.array_equal(sa[k], sb[k]), k
    assert np.abs(sc["Tt1o"] - sa["Tt1o"]).max() <= 1e-5
    assert np.abs(np.log(sc["slipv"] / sa["slipv"])).max() <= 1e-2


@pytest.mark.parametrize("accumulate", ["f64", "mfma"])
def test_device_ensemble_multi_iteration_graph_matches_eager(gpu, pq3d_lib, accumulate):
    """capture(steps=5): one graph of five iterations, one launch per five
    iterations; after every replay the state equals five eager iterations
    bit for bit, also while instances finish inside a replay (their later
    iterations skip them)."""
    from dataclasses import replace

    from pyquake3d_amd.hmatrix import HMatrix
    from pyquake3d_amd.qdsim import DeviceEnsemble

    base = planar_c1(nx=16, nz=8, seed=0)
    rng = np.random.default_rng(9)
    probs = [replace(base, a=base.a * (1 + 0.05 * rng.standard_normal(base.n)),
                     dc=base.dc * rng.uniform(0.8, 1.2)) for _ in range(16)]
    H = HMatrix.from_mesh(base.mesh, base.mu, base.lambda_, leaf=16, eta=2.0, eps=1e-6,
                          dtype=torch.float32, device=gpu)
    H.accumulate = accumulate
    # end times spread over the run (instance b ends where a probe run of
    # the same ensemble is after 8 + 6 b iterations), so that instances
    # finish inside replays
    probe = DeviceEnsemble(probs, H, device=gpu, rounds=4)
    times = []
    for _ in range(100):
        probe.step()
        times.append(probe.tdev.clone())
    t_end = torch.stack([times[7 + 6 * b][b] for b in range(16)])
    eager = DeviceEnsemble(probs, H, device=gpu, rounds=4)
    graph = DeviceEnsemble(probs, H, device=gpu, rounds=4)
    for e in (eager, graph):
        e.run[:, 0] = t_end
    fields = ("Tno", "Tt1o", "Tt2o", "state", "slipv", "slip1", "tdev", "hcur", "nacc", "flag")
    for _ in range(2):
        eager.step()
        graph.step()
    graph.capture(steps=5)
    with pytest.raises(ValueError):
        eager.capture(steps=0)
    for it in range(24):
        for _ in range(5):
            eager.step()
        graph.step()
        torch.cuda.synchronize()
        for f in fields:
            assert torch.equal(getattr(eager, f), getattr(graph, f)), (it, f)
    graph.sync()
    eager.sync()
    assert graph.iterations == eager.iterations == 2 + 24 * 5
    assert np.array_equal(graph.naccepted, eager.naccepted) and graph.naccepted.min() > 0
    assert graph.finished().all()


def test_ragged_ensemble_multi_iteration_graph_matches_eager(gpu, pq3d_lib):
    """The ragged ensemble (one block-diagonal operator over meshes of
    uneven size) as captured steps: graphs of four iterations replayed ten
    times give the state of forty eager iterations bit for bit."""
    from pyquake3d_amd.hmatrix import HMatrix, MultiHMatrix
    from pyquake3d_amd.problem import ensemble_c4
    from pyquake3d_amd.qdsim import DeviceEnsemble

    probs = ensemble_c4(8, seed=2, n_min=100, n_max=600)
    M = MultiHMatrix([HMatrix.from_mesh(p.mesh, p.mu, p.lambda_, leaf=16, eta=2.0, eps=1e-6,
                                        dtype=torch.float32, device=gpu) for p in probs])
    eager = DeviceEnsemble(probs, M, device=gpu, rounds=4)
    graph = DeviceEnsemble(probs, M, device=gpu, rounds=4)
    for _ in range(2):
        eager.step()
        graph.step()
    graph.capture(steps=4)
    fields = ("Tno", "Tt1o", "Tt2o", "state", "slipv", "slip1", "tdev", "hcur", "nacc", "flag")
    for it in range(10):
        for _ in range(4):
            eager.step()
        graph.step()
        torch.cuda.synchronize()
        for f in fields:
            assert torch.equal(getattr(eager, f), getattr(graph, f)), (it, f)
    graph.sync()
    eager.sync()
    assert graph.iterations == eager.iterations == 42
    assert np.array_equal(graph.naccepted, eager.naccepted) and graph.naccepted.min() > 0


@pytest.mark.parametrize("accumulate", ["f64", "mfma"])
def test_device_ensemble_graph_replay_matches_eager(gpu, pq3d_lib, accumulate):
    """A recorded step (HIP graph) replayed many times gives, after every
    replay, the state of the same number of eager steps bit for bit, with
    eager torch work and synchronisations between the replays; 16 draws
    (16-RHS matvec, FP64 sums or the matrix-core product) sharing one FP32
    operator. A new matrix-core configuration drops the graph."""
    from dataclasses import replace

    from pyquake3d_amd.hmatrix import HMatrix
    from pyquake3d_amd.qdsim import DeviceEnsemble

    base = planar_c1(nx=16, nz=8, seed=0)
    rng = np.random.default_rng(8)
    probs = [replace(base, a=base.a * (1 + 0.05 * rng.standard_normal(base.n)),
                     dc=base.dc * rng.uniform(0.8, 1.2)) for _ in range(16)]
    H = HMatrix.from_mesh(base.mesh, base.mu, base.lambda_, leaf=16, eta=2.0, eps=1e-6,
                          dtype=torch.float32, device=gpu)
    H.accumulate = accumulate
    eager = DeviceEnsemble(probs, H, device=gpu, rounds=4)
    graph = DeviceEnsemble(probs, H, device=gpu, rounds=4)
    fields = ("Tno", "Tt1o", "Tt2o", "state", "slipv", "slip1", "tdev", "hcur", "nacc", "flag")
    for _ in range(3):
        eager.step()
        graph.step()
    graph.capture()
    for it in range(40):
        eager.step()
        graph.step()
        junk = torch.randn(1000, device=gpu).sum()      # eager work between replays
        torch.cuda.synchronize()
        assert float(junk) == float(junk)
        for f in fields:
            assert torch.equal(getattr(eager, f), getattr(graph, f)), (it, f)
    graph.sync()
    eager.sync()
    assert graph.iterations == eager.iterations == 43
    assert np.array_equal(graph.naccepted, eager.naccepted) and graph.naccepted.min() > 0
    if accumulate == "mfma":
        H.tuning = {"depth": 2, "waves": 1, "order": "cost", "p1_depth": 2, "p1_waves": 2}
        for _ in range(3):
            eager.step()
            graph.step()
        assert graph._graph is None
        for f in fields:
            assert torch.equal(getattr(eager, f), getattr(graph, f)), ("retuned", f)
