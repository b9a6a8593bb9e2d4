"""Device controller run control, the two-plane product of a fixed normal
traction, and the learned step-size proposal on the device, against host
restatements of the same rules."""
from dataclasses import replace

import numpy as np
import pytest
import torch

from pyquake3d_amd import _lib
from pyquake3d_amd.problem import planar_c1, undulating_c3

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hm(gpu, pq3d_lib):
    from pyquake3d_amd.hmatrix import HMatrix
    p = undulating_c3(nx=40, nz=10, seed=3)
    out = {}
    for dt in (torch.float32, torch.float64):
        out[dt] = HMatrix.from_mesh(p.mesh, p.mu, p.lambda_, leaf=16, eta=2.0, eps=1e-6, dtype=dt,
                                    device=gpu)
    return p, out


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("nrhs", [1, 16])
def test_two_plane_product_equals_first_rows(gpu, hm, dtype, nrhs):
    """planes=2 (Fix_Tn: the Bs, Bd operators skipped, src/QDsim.py:906-909)
    gives the first two output rows of the full product bit for bit and
    leaves the third row untouched."""
    p, H = hm
    H = H[dtype]
    n = p.n
    g = torch.Generator(device="cpu").manual_seed(nrhs)
    v1 = torch.randn(n * nrhs, generator=g, dtype=torch.float64).to(gpu)
    v2 = torch.randn(n * nrhs, generator=g, dtype=torch.float64).to(gpu)
    full = H.matvec(v1, v2, nrhs=nrhs, instance_major=nrhs > 1)
    out = torch.full((3, n * nrhs), 7.0, dtype=torch.float64, device=gpu)
    H.matvec(v1, v2, out=out, nrhs=nrhs, instance_major=nrhs > 1, planes=2)
    assert torch.equal(out[:2], full[:2])
    assert bool((out[2] == 7.0).all())
    b3, b2 = H.algorithmic_bytes(nrhs, 3), H.algorithmic_bytes(nrhs, 2)
    assert b2 < b3 and b3 == H.matvec_bytes + 5 * 8 * n * (nrhs - 1)
    if dtype == torch.float32 and nrhs == 16:
        # the matrix-core product: the two-plane kernel computes the first two
        # planes exactly as the three-plane one
        H.accumulate = "mfma"
        try:
            full = H.matvec(v1, v2, nrhs=nrhs, instance_major=True)
            out.fill_(7.0)
            H.matvec(v1, v2, out=out, nrhs=nrhs, instance_major=True, planes=2)
        finally:
            H.accumulate = "f64"
        assert torch.equal(out[:2], full[:2])
        assert bool((out[2] == 7.0).all())


def test_fix_tn_steps_do_not_depend_on_the_skipped_rows(gpu, hm):
    """With a fixed normal traction the step reads only the loading for
    dsigma/dt, so stepping with the two-plane product equals stepping with
    the full product bit for bit."""
    from pyquake3d_amd.qdsim import DeviceEnsemble
    p, H = hm
    p = replace(p, fix_Tn=True)
    H = H[torch.float32]

    class Full:
        """The operator with the planes argument ignored (full product)."""
        def __init__(self, h):
            self.h, self.perm, self.n = h, h.perm, h.n
            self.accumulate = h.accumulate

        def matvec(self, v1, v2, out=None, nrhs=1, instance_major=False, planes=3):
            return self.h.matvec(v1, v2, out=out, nrhs=nrhs, instance_major=instance_major)

    a = DeviceEnsemble([p, p], H, device=gpu, rounds=4)
    b = DeviceEnsemble([p, p], Full(H), device=gpu, rounds=4)
    for _ in range(150):
        a.step()
        b.step()
    a.sync()
    b.sync()
    assert np.array_equal(a.time, b.time)
    assert torch.equal(a.slipv, b.slipv) and torch.equal(a.Tno, b.Tno)


def test_run_control_end_time_and_events(gpu, pq3d_lib):
    """Event start times and the end time kept on the device equal what a
    host loop reading max V after every step finds (same thresholds)."""
    from pyquake3d_amd.hmatrix import HMatrix
    from pyquake3d_amd.qdsim import DeviceEnsemble
    p = planar_c1(nx=12, nz=6, seed=0)
    H = HMatrix.from_mesh(p.mesh, p.mu, p.lambda_, leaf=16, eta=2.0, eps=1e-8,
                          dtype=torch.float64, device=gpu)
    year = 3.15576e7
    t_end = 260 * year
    vup, vdown = 1e-3, 1e-5
    dev = DeviceEnsemble([p], H, device=gpu, rounds=8, t_end=t_end, vup=vup, vdown=vdown)
    host = DeviceEnsemble([p], H, device=gpu, rounds=8)
    events, armed, peak, last = [], False, 0.0, 0
    while True:
        host.step()
        host.sync()
        if host.naccepted[0] == last:
            continue
        last = host.naccepted[0]
        v = float(host.max_slipv()[0])
        peak = max(peak, v)
        t = float(host.time[0])
        if armed and v > vup:
            events.append(t)
            armed = False
        elif v < vdown:
            armed = True
        if t >= t_end:
            break
    for _ in range(100000):
        dev.step()
        if dev.iterations % 200 == 0 and dev.finished()[0]:
            break
    dev.sync()
    assert dev.finished()[0]
    assert dev.naccepted[0] == host.naccepted[0]
    assert dev.time[0] == host.time[0]
    assert len(events) >= 1
    assert dev.events()[0].tolist() == events
    assert dev.peak_slipv()[0] == peak
    # a finished instance takes no more steps
    n0 = dev.naccepted[0]
    for _ in range(5):
        dev.step()
    dev.sync()
    assert dev.naccepted[0] == n0


def _trained_mlp(gpu, X, T):
    from pyquake3d_amd.learned_dt import StepMLP
    mlp = StepMLP(device=gpu, seed=0)
    mlp.fit(X, T, epochs=3, batch=256, lr=2e-3)
    return mlp


def test_device_proposal_matches_host_rule(gpu, hm):
    """pq3d_learned_propose: the features it records equal the host decode of
    pq3d_features, and its first-attempt step equals the host rule
    dtnext 10^min(prediction - margin, hi) with the host network forward."""
    from pyquake3d_amd.learned_dt import decode_features
    from pyquake3d_amd.qdsim import DeviceEnsemble
    p, H = hm
    H = H[torch.float32]
    draws = [p] + [undulating_c3(nx=40, nz=10, seed=3, friction_seed=k, mesh=p.mesh)
                   for k in range(1, 4)]
    rec = DeviceEnsemble(draws, H, device=gpu, rounds=8, record=4 * 400)
    for _ in range(300):
        rec.step()
    rec.sync()
    X, T = rec.samples()
    assert len(T) >= 4 * 250 and np.all(np.isfinite(X)) and np.all(T <= 1e-12)
    mlp = _trained_mlp(gpu, X, T)
    margin, hi = 0.05, 0.1
    ens = DeviceEnsemble(draws, H, device=gpu, rounds=8, mlp=mlp, margin=margin, hi=hi)
    for _ in range(40):
        ens.step()
    ens.sync()
    fresh = (ens.flag.cpu().numpy() == 0) & (ens.nrj.cpu().numpy() == 0)
    base = ens.hcur.cpu().numpy().copy()
    ls = ens.lstate.cpu().numpy().copy()
    _lib.call("pq3d_features", ens.ne, ens.nb, _lib.ptr(ens.inst), _lib.ptr(ens.slipv),
              _lib.ptr(ens.state), _lib.ptr(ens.b), _lib.ptr(ens.f0), ens.V0,
              _lib.ptr(ens.facc), _lib.stream_handle())
    acc = ens.facc.cpu().numpy().reshape(ens.nb, 4)
    _lib.call("pq3d_learned_propose", ens.nb, _lib.ptr(ens.facc), _lib.ptr(ens.ncell),
              _lib.ptr(ens.flag), _lib.ptr(ens.nrj), _lib.ptr(ens.hcur), _lib.ptr(ens.lstate),
              _lib.ptr(mlp.prm), _lib.ptr(ens.norm), _lib.stream_handle())
    got_h = ens.hcur.cpu().numpy()
    got_ls = ens.lstate.cpu().numpy()
    vmax, vmin, mean, frac = decode_features(acc, np.diff(ens.off))
    f = np.stack([vmax, vmin, mean, frac, np.log10(np.maximum(ls[:, 1], 1e-12)),
                  np.log10(ls[:, 2])], axis=1)
    pred = mlp.predict_log_dt(f)
    assert fresh.any()
    for b in range(ens.nb):
        if not fresh[b]:
            assert got_h[b] == base[b]
            continue
        assert np.allclose(got_ls[b, 4:10], f[b], rtol=1e-12, atol=1e-12)
        assert got_ls[b, 3] == base[b] and got_ls[b, 0] == 1.0
        want = base[b] * 10.0 ** min(pred[b] - margin, hi)
        assert got_h[b] == pytest.approx(want, rel=1e-5)


def test_commit_keeps_feature_slots_current(gpu, hm):
    """The feature slots of the learned proposal are reduced in full only
    once; afterwards the commit recomputes the accepted instances' slots from
    their new state. After every iteration the slots of the instances about
    to start a fresh step (the only ones the proposal reads) equal a full
    pq3d_features pass bit for bit, a run that reduces in full every step
    follows the same trajectory bit for bit, and a state load marks the
    slots stale."""
    from pyquake3d_amd.qdsim import DeviceEnsemble
    p, H = hm
    H = H[torch.float32]
    draws = [p] + [undulating_c3(nx=40, nz=10, seed=3, friction_seed=k, mesh=p.mesh)
                   for k in range(1, 4)]
    rec = DeviceEnsemble(draws, H, device=gpu, rounds=8, record=4 * 300)
    for _ in range(300):
        rec.step()
    rec.sync()
    X, T = rec.samples()
    mlp = _trained_mlp(gpu, X, T)
    # two attempt rounds per iteration: some instances carry a step over
    fused = DeviceEnsemble(draws, H, device=gpu, rounds=2, mlp=mlp)
    full = DeviceEnsemble(draws, H, device=gpu, rounds=2, mlp=mlp)
    tmp = torch.zeros_like(fused.facc)
    checked = carried = 0
    for _ in range(120):
        fused.step()
        full._feat_current = False          # full reduction at every step
        full.step()
        fresh = ((fused.flag == 0) & (fused.nrj == 0)).cpu().numpy()
        _lib.call("pq3d_features", fused.ne, fused.nb, _lib.ptr(fused.inst),
                  _lib.ptr(fused.slipv), _lib.ptr(fused.state), _lib.ptr(fused.b),
                  _lib.ptr(fused.f0), fused.V0, _lib.ptr(tmp), _lib.stream_handle())
        got = fused.facc.cpu().numpy().reshape(-1, 4)
        want = tmp.cpu().numpy().reshape(-1, 4)
        assert np.array_equal(got[fresh], want[fresh])
        checked += int(fresh.sum())
        carried += int((~fresh).sum())
    assert checked > 0 and carried > 0
    fused.sync()
    full.sync()
    assert not fused.aborted().any()
    for name in ("state", "Tt1o", "slipv", "slip1", "hcur", "tdev", "nacc", "lstate"):
        assert torch.equal(getattr(fused, name), getattr(full, name)), name
    d = fused.save_state()
    fused.load_state(d)
    assert not fused._feat_current
    fused.step()
    assert fused._feat_current


def test_learned_loop_on_device(gpu, hm):
    """The learned controller on the device end to end (reduced scale):
    trained on the reference controller's own rows, it takes every step
    through the estimator (no abort, every accepted error ratio below 1)
    and needs fewer attempts per accepted step than the reference
    controller over the same number of steps."""
    from pyquake3d_amd.qdsim import DeviceEnsemble
    p, H = hm
    H = H[torch.float32]
    draws = [p] + [undulating_c3(nx=40, nz=10, seed=3, friction_seed=k, mesh=p.mesh)
                   for k in range(1, 4)]
    nstep = 600
    rec = DeviceEnsemble(draws, H, device=gpu, rounds=1, record=4 * nstep)
    for _ in range(nstep):
        rec.step()
    rec.sync()
    X, T = rec.samples()
    mlp = _trained_mlp(gpu, X, T)
    plain = DeviceEnsemble(draws, H, device=gpu, rounds=1)
    learned = DeviceEnsemble(draws, H, device=gpu, rounds=1, mlp=mlp)
    for e in (plain, learned):
        for _ in range(nstep):
            e.step()
        e.sync()
    assert not learned.aborted().any()
    # one attempt per iteration: iterations / accepted steps = attempts per step
    a_plain = nstep * 4 / plain.naccepted.sum()
    a_learned = nstep * 4 / learned.naccepted.sum()
    assert a_learned < a_plain, (a_learned, a_plain)
    ls = learned.lstate.cpu().numpy()
    assert np.all(ls[:, 1] < 1.0)       # every accepted error ratio passed the test
