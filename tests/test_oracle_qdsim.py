"""Pin the CPU oracle of the time step (oracle/qdsim_oracle.py) against
mathematics that does not come from the reference's code: the published
Dormand-Prince 5(4) tableau and its order conditions, the embedded error's
order, the aging-law steady state, and the regularized rate-and-state law
differentiated numerically."""
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle import qdsim_oracle as Q

# Dormand & Prince (1980), "A family of embedded Runge-Kutta formulae"
A_PUB = {
    (2, 1): Fr(1, 5),
    (3, 1): Fr(3, 40), (3, 2): Fr(9, 40),
    (4, 1): Fr(44, 45), (4, 2): Fr(-56, 15), (4, 3): Fr(32, 9),
    (5, 1): Fr(19372, 6561), (5, 2): Fr(-25360, 2187), (5, 3): Fr(64448, 6561),
    (5, 4): Fr(-212, 729),
    (6, 1): Fr(9017, 3168), (6, 2): Fr(-355, 33), (6, 3): Fr(46732, 5247),
    (6, 4): Fr(49, 176), (6, 5): Fr(-5103, 18656),
}
B5_PUB = [Fr(35, 384), 0, Fr(500, 1113), Fr(125, 192), Fr(-2187, 6784), Fr(11, 84), 0]
B4_PUB = [Fr(5179, 57600), 0, Fr(7571, 16695), Fr(393, 640), Fr(-92097, 339200), Fr(187, 2100),
          Fr(1, 40)]


def test_tableau_matches_published_coefficients():
    got = {(2, 1): Q.B21, (3, 1): Q.B31, (3, 2): Q.B32, (4, 1): Q.B41, (4, 2): Q.B42,
           (4, 3): Q.B43, (5, 1): Q.B51, (5, 2): Q.B52, (5, 3): Q.B53, (5, 4): Q.B54,
           (6, 1): Q.B61, (6, 2): Q.B62, (6, 3): Q.B63, (6, 4): Q.B64, (6, 5): Q.B65}
    for k, v in A_PUB.items():
        if k == (6, 3):
            # the reference's B63 has the opposite sign (src/QDsim.py:1019)
            assert got[k] == -float(v)
        else:
            assert got[k] == float(v), k
    b5 = [Q.B71, 0, Q.B73, Q.B74, Q.B75, Q.B76, 0]
    b4 = [Q.B81, 0, Q.B83, Q.B84, Q.B85, Q.B86, Q.B87]
    assert b5 == [float(x) for x in B5_PUB]
    assert b4 == [float(x) for x in B4_PUB]


def test_order_conditions_exact():
    # c_i = row sums; stage 7 is the FSAL stage with a_7j = b5_j
    A = {i: {j: A_PUB.get((i, j), Fr(0)) for j in range(1, 8)} for i in range(1, 8)}
    for j in range(1, 8):
        A[7][j] = B5_PUB[j - 1]
    c = [sum(A[i].values()) for i in range(1, 8)]
    b5, b4 = B5_PUB, B4_PUB
    for b, order in ((b5, 5), (b4, 4)):
        for q in range(1, order + 1):  # bushy-tree conditions sum b c^(q-1) = 1/q
            assert sum(bi * ci ** (q - 1) for bi, ci in zip(b, c)) == Fr(1, q)
    # sum b_i a_ij c_j = 1/6 (third-order tall tree)
    for b in (b5, b4):
        s = sum(b[i - 1] * A[i][j] * c[j - 1] for i in range(1, 8) for j in range(1, 8))
        assert s == Fr(1, 6)


def _one_cell(a=0.01, b=0.015, dc=0.05, V=1e-6, tau_over=1.0, sigma=25.0, Kself=0.0):
    """Single-cell oracle: K is a 1x1 self-stiffness in A1s."""
    V0, mu, Cs, f0 = 1e-6, 32e9, 3464.0, 0.6
    Vpl = np.array([1e-9])
    psi = f0 + b * np.log(V0 / V)  # steady state at V
    Tt = sigma * a * np.arcsinh(V / (2 * V0) * np.exp(psi / a)) * tau_over
    z = np.zeros((1, 1))
    K = dict(A1s=np.array([[Kself]]), A2s=z, Bs=z, A1d=z, A2d=z, Bd=z)
    init = (np.array([sigma]), np.array([Tt]), np.array([0.0]), np.array([psi]),
            np.array([V]), np.array([1e-16]), np.array([V]))
    sim = Q.OracleSim(K, np.array([a]), np.array([b]), np.array([dc]), np.array([f0]), V0, mu,
                      Cs, Vpl, np.array([0.0]), init)
    return sim


def test_steady_state_is_fixed_point_of_state_law():
    sim = _one_cell(V=3e-7)
    sim.AdotV1 = np.zeros(1)
    sim.AdotV2 = np.zeros(1)
    sim.dsigmadt = np.zeros(1)
    dpsi, dsig, dt1, dt2 = sim.derivative(sim.Tno, sim.Tt1o, sim.Tt2o, sim.state)
    assert abs(dpsi[0]) < 1e-12 * (sim.b[0] / sim.dc[0]) * 3e-7 * 1e3
    assert dt1[0] == 0.0 or abs(dt1[0]) < 1e-15


def test_traction_rate_matches_numerical_derivatives_of_the_law():
    """dtau/dt = (-AdotV + load - eta (dV/dsigma sdot + dV/dpsi psidot)) /
    (1 + eta dV/dtau), with V = 2 V0 exp(-psi/a) sinh(tau/(a sigma)) --
    derivatives taken by central differences, independent of the closed
    forms in the oracle."""
    sim = _one_cell(V=2e-8, tau_over=1.02)
    sim.AdotV1 = np.array([3.7e3])
    sim.AdotV2 = np.zeros(1)
    sim.dsigmadt = np.array([-1.3e2])
    sim.slipv = np.array([2.5e-8])
    a, b, dc, f0, V0 = sim.a[0], sim.b[0], sim.dc[0], sim.f0[0], sim.V0
    s, t, p = sim.Tno[0] * 1e6, sim.Tt1o[0] * 1e6, sim.state[0]

    def V(t, s, p):
        return 2 * V0 * np.exp(-p / a) * np.sinh(t / (a * s))

    dVdt = (V(t * (1 + 1e-7), s, p) - V(t * (1 - 1e-7), s, p)) / (2e-7 * t)
    dVds = (V(t, s * (1 + 1e-7), p) - V(t, s * (1 - 1e-7), p)) / (2e-7 * s)
    dVdp = (V(t, s, p + 1e-7) - V(t, s, p - 1e-7)) / 2e-7
    psidot = b / dc * (V0 * np.exp((f0 - p) / b) - sim.slipv[0])
    eta = sim.mu / (2 * sim.Cs)
    expect = (-sim.AdotV1[0] - eta * (dVds * sim.dsigmadt[0] + dVdp * psidot)) / (1 + eta * dVdt)
    dpsi, dsig, dt1, dt2 = sim.derivative(sim.Tno, sim.Tt1o, sim.Tt2o, sim.state)
    assert dpsi[0] == pytest.approx(psidot, rel=1e-12)
    assert dt1[0] * 1e6 == pytest.approx(expect, rel=1e-6)
    assert dsig[0] * 1e6 == pytest.approx(sim.dsigmadt[0], rel=1e-15)


def _err_ratios(tableau, hs):
    sim = _one_cell(V=1e-7, tau_over=1.01, Kself=3e7)
    sim.b63 = Q.B63 if tableau == "ref" else -Q.B63
    sim.simu_forward(1e-3)  # sets the frozen AdotV etc.
    errs = []
    for h in hs:
        _, c1, c2, _, _ = sim.rk45_attempt(h)
        errs.append(c1 * sim.reltol - 1e-10)  # psi error (largest here)
    return errs[0] / errs[1], errs[1] / errs[2]


def test_embedded_error_is_fifth_order_with_published_tableau():
    r1, r2 = _err_ratios("dp5", (1e5, 5e4, 2.5e4))
    # local error of the 4th-order member ~ h^5 -> ratio 32 per halving
    assert 28 < r1 < 40 and 28 < r2 < 40, (r1, r2)


def test_reference_tableau_error_is_second_order():
    """The reference's sign of B63 makes stage 6 inconsistent (row sum
    -16.8): the embedded error then falls only like h^2."""
    r1, r2 = _err_ratios("ref", (1e3, 5e2, 2.5e2))
    assert 3.5 < r1 < 4.5 and 3.5 < r2 < 4.5, (r1, r2)


def test_controller_rejects_and_shrinks_on_large_step():
    sim = _one_cell(V=1e-7, tau_over=1.05, Kself=3e7)
    h, dtn = sim.simu_forward(1e6)
    assert h < 1e6          # the huge trial step was rejected at least once
    assert dtn <= 1.5 * h
