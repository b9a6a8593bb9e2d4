"""CPU oracle of the half-space TDE stress -- TEST INFRASTRUCTURE ONLY.

Only tests/ use it, as the checker. Half-space stress at receivers z <= 0
(src/TDstressFS_C.cpp:1911-1999, TDstressHS):

    S = S_main + S_harmonic + S_image

  * S_main  : full-space stress of the triangle (the C oracle, tde_oracle.c)
  * S_image : full-space stress of the triangle mirrored in z = 0; its xz and
              yz components change sign when all three vertices lie at z = 0
              (:1954-1963)
  * S_harm  : the free-surface correction of the three angular dislocations
              along the sides (TDstress_HarFunc :1766-1908, AngSetupFSC_S
              :1489-1763). Their strain is computed here as the symmetric
              gradient of the free-surface-correction displacement of one
              angular dislocation (Nikkhoo & Walter 2015; the reference's own
              displacement twin src/DH_greenfunction.py:313-399), with
              complex-step derivatives (exact to rounding), instead of the
              reference's expanded strain formula (:613-1486).

The reference skips a side that is vertical (beta within 2.2204e-16 of 0 or
pi, :1524-1531) without writing its outputs, so TDstress_HarFunc adds the
previous side's values again (zero for the first side). ``stale=True``
reproduces that; ``stale=False`` gives the side a zero contribution, as the
published method does.

Pinned by tests/test_oracle_tde_hs.py: golden vectors of the reference C++
(compiled from source, tests/golden/g4_halfspace.npz) and the traction-free
free surface (sigma_xz = sigma_yz = sigma_zz = 0 at z = 0) with stale=False.
"""
from __future__ import annotations

import numpy as np

PI = 3.141592653589793
EPS_VERT = 2.2204e-16   # src/TDstressFS_C.cpp:1524


def fsc_displacement(y1, y2, y3, sinB, cosB, cotB, tanH, b, nu, a):
    """Free-surface correction displacement of one angular dislocation at
    (y1, y2, y3) in its frame (src/DH_greenfunction.py:313-399); complex
    inputs allowed (complex-step differentiation)."""
    b1, b2, b3 = b
    N1 = 1 - 2 * nu
    M1 = 1 - nu
    y3b = y3 + 2 * a
    z1b = y1 * cosB + y3b * sinB
    z3b = -y1 * sinB + y3b * cosB
    r2b = y1 * y1 + y2 * y2 + y3b * y3b
    rb = np.sqrt(r2b)
    r3b = rb * r2b
    W6 = rb + y3b
    W7 = rb + z3b
    Fib = 2 * np.arctan(-y2 / (-W6 / tanH + y1))
    L6, L7 = np.log(W6), np.log(W7)
    ar = a / rb
    h = y3b - a
    q = rb * cosB + y3b
    c2 = cotB * cotB
    u1 = (-2 * M1 * N1 * Fib * c2 + N1 * y2 / W6 * ((N1 - ar) * cotB - y1 / W6 * (nu + ar))
          + N1 * y2 * cosB * cotB / W7 * (cosB + ar) + a * y2 * h * cotB / r3b
          + y2 * h / (rb * W6) * (-N1 * cotB + y1 / W6 * (2 * nu + ar) + a * y1 / r2b)
          + y2 * h / (rb * W7) * (cosB / W7 * (q * (N1 * cosB - ar) * cotB
                                             + 2 * M1 * (rb * sinB - y1) * cosB)
                                  - a * y3b * cosB * cotB / r2b))
    u2 = (N1 * ((2 * M1 * c2 - nu) * L6 - (2 * M1 * c2 + N1) * cosB * L7)
          - N1 / W6 * (y1 * cotB * (N1 - ar) + nu * y3b - a + y2 * y2 / W6 * (nu + ar))
          - N1 * z1b * cotB / W7 * (cosB + ar) - a * y1 * h * cotB / r3b
          + h / W6 * (-2 * nu + 1 / rb * (N1 * y1 * cotB - a) + y2 * y2 / (rb * W6) * (2 * nu + ar)
                      + a * y2 * y2 / r3b)
          + h / W7 * (cosB * cosB - 1 / rb * (N1 * z1b * cotB + a * cosB) + a * y3b * z1b * cotB / r3b
                      - 1 / (rb * W7) * (y2 * y2 * cosB * cosB - a * z1b * cotB / rb * q)))
    u3 = (2 * M1 * (N1 * Fib * cotB + y2 / W6 * (2 * nu + ar) - y2 * cosB / W7 * (cosB + ar))
          + y2 * h / rb * (2 * nu / W6 + a / r2b)
          + y2 * h * cosB / (rb * W7) * (N1 - q / W7 * (cosB + ar) - a * y3b / r2b))
    w1 = (N1 * ((2 * M1 * c2 + nu) * L6 - (2 * M1 * c2 + 1) * cosB * L7)
          + N1 / W6 * (-N1 * y1 * cotB + nu * y3b - a + a * y1 * cotB / rb + y1 * y1 / W6 * (nu + ar))
          - N1 * cotB / W7 * (z1b * cosB - a * (rb * sinB - y1) / (rb * cosB)) - a * y1 * h * cotB / r3b
          + h / W6 * (2 * nu + 1 / rb * (N1 * y1 * cotB + a) - y1 * y1 / (rb * W6) * (2 * nu + ar)
                      - a * y1 * y1 / r3b)
          + h * cotB / W7 * (-cosB * sinB + a * y1 * y3b / (r3b * cosB)
                             + (rb * sinB - y1) / rb * (2 * M1 * cosB - q / W7 * (1 + a / (rb * cosB)))))
    w2 = (2 * M1 * N1 * Fib * c2 + N1 * y2 / W6 * (-(N1 - ar) * cotB + y1 / W6 * (nu + ar))
          - N1 * y2 * cotB / W7 * (1 + a / (rb * cosB)) - a * y2 * h * cotB / r3b
          + y2 * h / (rb * W6) * (N1 * cotB - 2 * nu * y1 / W6 - a * y1 / rb * (1 / rb + 1 / W6))
          + y2 * h * cotB / (rb * W7) * (-2 * M1 * cosB + q / W7 * (1 + a / (rb * cosB))
                                         + a * y3b / (r2b * cosB)))
    w3 = (-2 * M1 * N1 * cotB * (L6 - cosB * L7) - 2 * M1 * y1 / W6 * (2 * nu + ar)
          + 2 * M1 * z1b / W7 * (cosB + ar) + h / rb * (N1 * cotB - 2 * nu * y1 / W6 - a * y1 / r2b)
          - h / W7 * (cosB * sinB + q * cotB / rb * (2 * M1 * cosB - q / W7)
                      + ar * (sinB - y3b * z1b / r2b - z1b * q / (rb * W7))))
    s1 = (N1 * (y2 / W6 * (1 + ar) - y2 * cosB / W7 * (cosB + ar)) - y2 * h / rb * (a / r2b + 1 / W6)
          + y2 * h * cosB / (rb * W7) * (q / W7 * (cosB + ar) + a * y3b / r2b))
    s2 = (N1 * (-sinB * L7 - y1 / W6 * (1 + ar) + z1b / W7 * (cosB + ar))
          + y1 * h / rb * (a / r2b + 1 / W6)
          - h / W7 * (sinB * (cosB - ar) + z1b / rb * (1 + a * y3b / r2b)
                      - 1 / (rb * W7) * (y2 * y2 * cosB * sinB - a * z1b / rb * q)))
    s3 = (2 * M1 * Fib + 2 * M1 * (y2 * sinB / W7 * (cosB + ar))
          + y2 * h * sinB / (rb * W7) * (1 + q / W7 * (cosB + ar) + a * y3b / r2b))
    c = 1.0 / 4.0 / PI / (1 - nu)
    return (c * (b1 * u1 + b2 * w1 + b3 * s1), c * (b1 * u2 + b2 * w2 + b3 * s2),
            c * (b1 * u3 + b2 * w3 + b3 * s3))


def fsc_strain(y, beta, b, nu, a):
    """Strain (xx, yy, zz, xy, xz, yz) in the angular dislocation's frame:
    symmetric gradient of fsc_displacement by complex steps."""
    trig = (np.sin(beta), np.cos(beta), 1.0 / np.tan(beta), np.tan(beta / 2.0))
    G = np.zeros((len(y), 3, 3))
    h = 1e-30
    for j in range(3):
        yc = y.astype(complex)
        yc[:, j] += 1j * h
        u = fsc_displacement(yc[:, 0], yc[:, 1], yc[:, 2], *trig, b, nu, a)
        for i in range(3):
            G[:, i, j] = np.imag(u[i]) / h
    E = 0.5 * (G + np.transpose(G, (0, 2, 1)))
    return np.stack([E[:, 0, 0], E[:, 1, 1], E[:, 2, 2], E[:, 0, 1], E[:, 0, 2], E[:, 1, 2]], 1)


def _rotate(E, A):
    """Local -> global, T' = A T A^T with A columns (ey1, ey2, ey3)."""
    T = np.zeros((len(E), 3, 3))
    idx = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
    for k, (i, j) in enumerate(idx):
        T[:, i, j] = E[:, k]
        T[:, j, i] = E[:, k]
    Tg = A @ T @ A.T
    return np.stack([Tg[:, i, j] for i, j in idx], 1)


def _hooke(E, mu, lam):
    tr = E[:, 0] + E[:, 1] + E[:, 2]
    return np.concatenate([2 * mu * E[:, :3] + lam * tr[:, None], 2 * mu * E[:, 3:]], 1)


def harmonic(pts, P, Ss, Ds, Ts, mu, lam, stale=True):
    """Free-surface correction stress and strain (n x 6 each) of triangle P
    (3 x 3) with Burgers (Ss, Ds, Ts) at points pts (n x 3)
    (src/TDstressFS_C.cpp:1766-1908)."""
    pts = np.asarray(pts, dtype=np.float64)
    P = np.asarray(P, dtype=np.float64)
    nu = 1.0 / (1.0 + lam / mu) / 2.0
    Vn = np.cross(P[1] - P[0], P[2] - P[0])
    if np.linalg.norm(Vn) != 0:
        Vn = Vn / np.linalg.norm(Vn)
    Vs = np.cross([0.0, 0.0, 1.0], Vn)
    if np.linalg.norm(Vs) == 0:
        Vs = np.array([0.0, 1.0, 0.0]) * Vn[2]
    else:
        Vs = Vs / np.linalg.norm(Vs)
    Vd = np.cross(Vn, Vs)
    B = Vn * Ts + Vs * Ss + Vd * Ds
    S = np.zeros((len(pts), 6))
    E = np.zeros((len(pts), 6))
    last = None
    for PA, PB in ((P[0], P[1]), (P[1], P[2]), (P[2], P[0])):
        side = PB - PA
        beta = np.arccos(-side[2] / np.linalg.norm(side))
        if abs(beta) < EPS_VERT or abs(PI - beta) < EPS_VERT:
            if stale and last is not None:
                S = S + last[0]
                E = E + last[1]
            continue
        ey1 = np.array([side[0], side[1], 0.0])
        ey1 = ey1 / np.linalg.norm(ey1[:2])
        ey3 = np.array([0.0, 0.0, -1.0])
        ey2 = np.cross(ey3, ey1)
        A = np.stack([ey1, ey2, ey3], 1)
        yA = (pts - PA) @ A
        yB = yA - side @ A
        bv = A.T @ B
        conf = beta * yA[:, 0] >= 0
        eA = np.zeros((len(pts), 6))
        eB = np.zeros((len(pts), 6))
        if conf.any():
            fl = np.array([-1.0, -1.0, 1.0])
            a1 = fsc_strain(yA[conf] * fl, PI - beta, bv * fl, nu, -PA[2])
            b1 = fsc_strain(yB[conf] * fl, PI - beta, bv * fl, nu, -PB[2])
            a1[:, 4:6] *= -1
            b1[:, 4:6] *= -1
            eA[conf], eB[conf] = a1, b1
        if (~conf).any():
            eA[~conf] = fsc_strain(yA[~conf], beta, bv, nu, -PA[2])
            eB[~conf] = fsc_strain(yB[~conf], beta, bv, nu, -PB[2])
        Ee = _rotate(eB - eA, A)
        Se = _hooke(Ee, mu, lam)
        last = (Se, Ee)
        S = S + Se
        E = E + Ee
    return S, E


def stress_hs(fs_oracle, pts, P, Ss, Ds, Ts, mu, lam, stale=True):
    """Half-space stress and strain: main + harmonic + image
    (src/TDstressFS_C.cpp:1911-1999). ``fs_oracle`` is oracle.tde_oracle.Oracle."""
    pts = np.asarray(pts, dtype=np.float64)
    P = np.asarray(P, dtype=np.float64)
    sm, em = fs_oracle.stress_points(pts, P[0], P[1], P[2], Ss, Ds, Ts, mu, lam)
    sh, eh = harmonic(pts, P, Ss, Ds, Ts, mu, lam, stale=stale)
    Pi = P.copy()
    Pi[:, 2] = -Pi[:, 2]
    si, ei = fs_oracle.stress_points(pts, Pi[0], Pi[1], Pi[2], Ss, Ds, Ts, mu, lam)
    if P[0, 2] == 0 and P[1, 2] == 0 and P[2, 2] == 0:
        si = si.copy()
        ei = ei.copy()
        si[:, 4:6] *= -1
        ei[:, 4:6] *= -1
    return sm + sh + si, em + eh + ei
