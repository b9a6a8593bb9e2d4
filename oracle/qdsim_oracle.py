"""CPU oracle of the time step -- TEST INFRASTRUCTURE ONLY.

Only tests/ and the CPU-baseline leg of bench.py use it, as the checker.
A numpy restatement of the reference's step on a dense FP64 operator:

  derivative        follows src/QDsim.py:760-813
  rk45_attempt      follows src/QDsim.py:1001-1134 (Dormand-Prince 5(4),
                    max-relative errors divided by the 4th-order solution)
  simu_forward      follows src/QDsim.py:892-996 (deficit, one frozen operator
                    product per step, accept/reject controller, state update)

Pinned by tests/test_oracle_qdsim.py: the tableau against the published
Dormand-Prince coefficients and order conditions, the embedded pair's error
order on a scalar test ODE, the steady state as a fixed point of the state
law, and single-cell relations of the regularized law.
"""
from __future__ import annotations

import numpy as np

# Dormand-Prince 5(4) as written in src/QDsim.py:1002-1032. B63 there is
# -46732/5247; the published coefficient is +46732/5247 (row 6 then sums to 1).
# The oracle keeps the reference's value by default ("ref") and offers the
# published one ("dp5").
B21 = .2
B31, B32 = 3. / 40, 9. / 40.
B41, B42, B43 = 44. / 45., -56. / 15, 32. / 9
B51, B52, B53, B54 = 19372. / 6561., -25360 / 2187., 64448. / 6561., -212. / 729.
B61, B62, B63, B64, B65 = 9017. / 3168., -355. / 33., -46732. / 5247., 49. / 176., -5103. / 18656.
B71, B73, B74, B75, B76 = 35. / 384., 500. / 1113., 125. / 192., -2187. / 6784., 11. / 84.
B81, B83, B84, B85, B86, B87 = (5179. / 57600., 7571. / 16695., 393. / 640., -92097. / 339200.,
                                187. / 2100., 1. / 40.)


class OracleSim:
    """Dense FP64 restatement of the reference's QDsim step.

    K: dict with the six (N, N) operators A1s, A2s, Bs, A1d, A2d, Bd.
    Arrays in original element order; tractions in MPa.
    """

    def __init__(self, K, a, b, dc, f0, V0, mu, Cs, Vpl, rake0, init, shear_loading=None,
                 normal_loading=None, fix_Tn=False, reltol=1e-4, tableau="ref", matvec=None):
        """K: dict of dense operators, or None with ``matvec(v1, v2) ->
        (AdotV1, AdotV2, dsigma_K)`` (e.g. a CPU H-matrix)."""
        self.K = K
        self.matvec = matvec
        self.a, self.b, self.dc, self.f0 = a, b, dc, f0
        self.V0, self.mu, self.Cs = V0, mu, Cs
        self.slipvC = Vpl
        self.rake0 = rake0
        n = len(a)
        self.shear_loading = np.zeros(n) if shear_loading is None else shear_loading
        self.normal_loading = np.zeros(n) if normal_loading is None else normal_loading
        self.fix_Tn = fix_Tn
        self.reltol = reltol
        self.b63 = B63 if tableau == "ref" else -B63
        (self.Tno, self.Tt1o, self.Tt2o, self.state, self.slipv1, self.slipv2,
         self.slipv) = [np.array(x, dtype=np.float64) for x in init]
        self.rake = np.array(rake0, dtype=np.float64)
        self.slip1 = np.zeros(n)
        self.slip2 = np.zeros(n)
        self.time = 0.0
        self.maxslipv0 = np.max(self.slipv)

    def derivative(self, Tno, Tt1o, Tt2o, state):
        Tno = Tno * 1e6
        Tt1o = Tt1o * 1e6
        Tt2o = Tt2o * 1e6
        V0, a, b, dc, f0 = self.V0, self.a, self.b, self.dc, self.f0
        e1p = np.exp(-state / a + Tt1o / (a * Tno))
        e1m = np.exp(-state / a - Tt1o / (a * Tno))
        e2p = np.exp(-state / a + Tt2o / (a * Tno))
        e2m = np.exp(-state / a - Tt2o / (a * Tno))
        dV1dtau = 2 * V0 / (a * Tno) * 0.5 * (e1p + e1m)
        dV2dtau = 2 * V0 / (a * Tno) * 0.5 * (e2p + e2m)
        dV1dsigma = -2 * V0 * Tt1o / (a * Tno ** 2) * 0.5 * (e1p + e1m)
        dV2dsigma = -2 * V0 * Tt2o / (a * Tno ** 2) * 0.5 * (e2p + e2m)
        dV1dstate = -2 * V0 / a * 0.5 * (e1p - e1m)
        dV2dstate = -2 * V0 / a * 0.5 * (e2p - e2m)
        dstatedt = b / dc * (V0 * np.exp(np.clip((f0 - state) / b, -700, 700)) - self.slipv)
        eta = self.mu / (2.0 * self.Cs)
        dtau1dt = (-self.AdotV1 + self.shear_loading - eta * (dV1dsigma * self.dsigmadt + dV1dstate * dstatedt)) / (1.0 + eta * dV1dtau)
        dtau2dt = (-self.AdotV2 + self.shear_loading - eta * (dV2dsigma * self.dsigmadt + dV2dstate * dstatedt)) / (1.0 + eta * dV2dtau)
        return dstatedt, self.dsigmadt / 1e6, dtau1dt / 1e6, dtau2dt / 1e6

    def rk45_attempt(self, h):
        Tno, Tt1o, Tt2o, state = self.Tno, self.Tt1o, self.Tt2o, self.state
        B63 = self.b63
        d = self.derivative
        p1, s1, t11, t21 = d(Tno, Tt1o, Tt2o, state)
        p2, s2, t12, t22 = d(Tno + h * B21 * s1, Tt1o + h * B21 * t11, Tt2o + h * B21 * t21,
                             state + h * B21 * p1)
        p3, s3, t13, t23 = d(Tno + h * (B31 * s1 + B32 * s2), Tt1o + h * (B31 * t11 + B32 * t12),
                             Tt2o + h * (B31 * t21 + B32 * t22), state + h * (B31 * p1 + B32 * p2))
        p4, s4, t14, t24 = d(Tno + h * (B41 * s1 + B42 * s2 + B43 * s3),
                             Tt1o + h * (B41 * t11 + B42 * t12 + B43 * t13),
                             Tt2o + h * (B41 * t21 + B42 * t22 + B43 * t23),
                             state + h * (B41 * p1 + B42 * p2 + B43 * p3))
        p5, s5, t15, t25 = d(Tno + h * (B51 * s1 + B52 * s2 + B53 * s3 + B54 * s4),
                             Tt1o + h * (B51 * t11 + B52 * t12 + B53 * t13 + B54 * t14),
                             Tt2o + h * (B51 * t21 + B52 * t22 + B53 * t23 + B54 * t24),
                             state + h * (B51 * p1 + B52 * p2 + B53 * p3 + B54 * p4))
        p6, s6, t16, t26 = d(Tno + h * (B61 * s1 + B62 * s2 + B63 * s3 + B64 * s4 + B65 * s5),
                             Tt1o + h * (B61 * t11 + B62 * t12 + B63 * t13 + B64 * t14 + B65 * t15),
                             Tt2o + h * (B61 * t21 + B62 * t22 + B63 * t23 + B64 * t24 + B65 * t25),
                             state + h * (B61 * p1 + B62 * p2 + B63 * p3 + B64 * p4 + B65 * p5))
        s7 = Tno + h * (B71 * s1 + B73 * s3 + B74 * s4 + B75 * s5 + B76 * s6)
        t17 = Tt1o + h * (B71 * t11 + B73 * t13 + B74 * t14 + B75 * t15 + B76 * t16)
        t27 = Tt2o + h * (B71 * t21 + B73 * t23 + B74 * t24 + B75 * t25 + B76 * t26)
        q7 = state + h * (B71 * p1 + B73 * p3 + B74 * p4 + B75 * p5 + B76 * p6)
        p7_, s7_, t17_, t27_ = d(s7, t17, t27, q7)
        s8 = Tno + h * (B81 * s1 + B83 * s3 + B84 * s4 + B85 * s5 + B86 * s6 + B87 * s7_)
        t18 = Tt1o + h * (B81 * t11 + B83 * t13 + B84 * t14 + B85 * t15 + B86 * t16 + B87 * t17_)
        t28 = Tt2o + h * (B81 * t21 + B83 * t23 + B84 * t24 + B85 * t25 + B86 * t26 + B87 * t27_)
        q8 = state + h * (B81 * p1 + B83 * p3 + B84 * p4 + B85 * p5 + B86 * p6 + B87 * p7_)
        with np.errstate(divide="ignore", invalid="ignore"):
            R1 = np.max(np.abs(np.abs(q8 - q7) / q8)) + 1e-10
            R2 = max(np.max(np.abs(np.abs(t18 - t17) / t18)), np.max(np.abs(np.abs(t28 - t27) / t28)),
                     np.max(np.abs(np.abs(s8 - s7) / s8))) + 1e-10
        c1 = R1 / self.reltol
        c2 = R2 / self.reltol
        hn1 = h * 0.9 * (self.reltol / R1) ** 0.2
        hn2 = h * 0.9 * (self.reltol / R2) ** 0.2
        return (s7, t17, t27, q7), c1, c2, hn1, hn2

    def simu_forward(self, dttry):
        K = self.K
        v1 = self.slipv1 - self.slipvC * np.cos(self.rake0)
        v2 = self.slipv2 - self.slipvC * np.sin(self.rake0)
        if self.matvec is not None:
            y1, y2, y3 = self.matvec(v1, v2)
            self.dsigmadt = self.normal_loading if self.fix_Tn else y3 + self.normal_loading
            self.AdotV1, self.AdotV2 = y1, y2
        else:
            if self.fix_Tn:
                self.dsigmadt = self.normal_loading
            else:
                self.dsigmadt = K["Bs"] @ v1 + K["Bd"] @ v2 + self.normal_loading
            self.AdotV1 = K["A1s"] @ v1 + K["A1d"] @ v2
            self.AdotV2 = K["A2s"] @ v1 + K["A2d"] @ v2
        nrjct = 0
        h = dttry
        while True:
            y, c1, c2, hn1, hn2 = self.rk45_attempt(h)
            if max(c1, c2) < 1.0 and not (np.isnan(c1) or np.isnan(c2)):
                dtnext = min(min(hn1, hn2), 1.5 * h)
                break
            nrjct += 1
            h = max(0.5 * h, min(hn1, hn2))
            if h < 1.e-15 or nrjct > 20:
                raise RuntimeError("dt is too small")
        self.time += h
        self.Tno, self.Tt1o, self.Tt2o, self.state = y
        self.slipv1 = (2.0 * self.V0) * np.exp(-self.state / self.a) * np.sinh(self.Tt1o / self.Tno / self.a)
        self.slipv2 = (2.0 * self.V0) * np.exp(-self.state / self.a) * np.sinh(self.Tt2o / self.Tno / self.a)
        self.slipv = np.sqrt(self.slipv1 * self.slipv1 + self.slipv2 * self.slipv2)
        self.rake = np.arctan2(self.Tt2o, self.Tt1o)
        self.slipv[self.slipv < 1e-30] = 1e-30
        self.maxslipv0 = np.max(self.slipv)
        self.slip1 = self.slip1 + self.slipv1 * h
        self.slip2 = self.slip2 + self.slipv2 * h
        return h, dtnext


def dense_operators(oracle, xg, eleVec, tri, mu, lam):
    """The six FP64 operators from the TDE oracle."""
    n = len(xg)
    K6 = oracle.traction_block(xg, eleVec, tri, np.arange(n), np.arange(n), mu, lam)
    return dict(zip(("A1s", "A2s", "Bs", "A1d", "A2d", "Bd"), K6))
