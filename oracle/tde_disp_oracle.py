"""CPU oracle of the TDE displacement -- TEST INFRASTRUCTURE ONLY.

Only tests/ use it, as the checker. Displacement of a triangular dislocation
(Nikkhoo & Walter 2015), as the reference's src/DH_greenfunction.py has it:

  td_disp_fs   TDdispFS (:99-220): three angular dislocations (AngDisDisp
               :65-96, TDSetupD :34-62) chosen by the configuration of the
               point (trimodefinder :6-31; points on a side in the TD plane
               give NaN), plus the Burgers-function term Fi times the slip.
  td_disp_hs   TDdispHS (:402-443): main + free-surface correction
               (TDdisp_HarFunc :223-258, AngSetupFSC :261-310, the
               displacement of oracle.tde_hs_oracle.fsc_displacement,
               AngDisDispFSC :313-399) + image; for a triangle with all
               vertices at z = 0 the image's vertical component and then
               the whole sum change sign.

A side whose direction is vertical contributes nothing to the correction,
as in the published method. (The reference's Python leaves that branch of
AngSetupFSC without a return value, so its TDdisp_HarFunc fails on such a
side; every triangle of a vertical ruled fault has one.)

Pinned by tests/test_oracle_tde_disp.py: the symmetric gradient of the
displacement (central differences) against the strain of the stress oracles
(tde_oracle.c, pinned to the reference C++ golden vectors; tde_hs_oracle
with a zero contribution of vertical sides), the jump of the displacement
across the triangle (equal to the Burgers vector), and far-field decay.
"""
from __future__ import annotations

import numpy as np

from .tde_hs_oracle import fsc_displacement

PI = 3.141592653589793


def _frame(P1, P2, P3):
    vn = np.cross(P2 - P1, P3 - P1)
    vn = vn / np.linalg.norm(vn)
    vs = np.cross(np.array([0.0, 0.0, 1.0]), vn)
    vs = vs / np.linalg.norm(vs)
    vd = np.cross(vn, vs)
    return vn, vs, vd


def _mode(q, p1, p2, p3):
    """+1 / -1 / 0 configuration of TD-coordinate points (normal, strike,
    dip) with the exact comparisons of the reference."""
    xp, yp, zp = q[:, 1], q[:, 2], q[:, 0]
    x1, x2, x3 = p1[1], p2[1], p3[1]
    y1, y2, y3 = p1[2], p2[2], p3[2]
    den = (x1 - x3) * (y2 - y3) - (x2 - x3) * (y1 - y3)
    a = ((xp - x3) * (y2 - y3) - (x2 - x3) * (yp - y3)) / den
    b = ((x1 - x3) * (yp - y3) - (xp - x3) * (y1 - y3)) / den
    c = 1.0 - a - b
    m = np.ones(len(xp))
    m[(a <= 0) & (b > c) & (c > a)] = -1
    m[(b <= 0) & (c > a) & (a > b)] = -1
    m[(c <= 0) & (a > b) & (b > c)] = -1
    m[(a == 0) & (b >= 0) & (c >= 0)] = 0
    m[(a >= 0) & (b == 0) & (c >= 0)] = 0
    m[(a >= 0) & (b >= 0) & (c == 0)] = 0
    m[(m == 0) & (zp != 0)] = 1
    return m


def ang_dis_disp(x, y, z, alpha, bx, by, bz, nu):
    """Displacement of one angular dislocation in its frame (AngDisDisp)."""
    cA, sA = np.cos(alpha), np.sin(alpha)
    eta = y * cA - z * sA
    zeta = y * sA + z * cA
    r = np.sqrt(x * x + y * y + z * z)
    zeta = np.where(zeta > r, r, zeta)     # keep the logarithms real
    z = np.where(z > r, r, z)
    k = 1.0 / 8.0 / PI / (1.0 - nu)
    N = 1.0 - 2.0 * nu
    rz, rzeta = r - z, r - zeta
    Lz, Lzeta = np.log(rz), np.log(rzeta)
    ux = bx * k * (x * y / r / rz - x * eta / r / rzeta)
    vx = bx * k * (eta * sA / rzeta - y * eta / r / rzeta + y * y / r / rz
                   + N * (cA * Lzeta - Lz))
    wx = bx * k * (eta * cA / rzeta - y / r - eta * z / r / rzeta - N * sA * Lzeta)
    uy = by * k * (x * x * cA / r / rzeta - x * x / r / rz - N * (cA * Lzeta - Lz))
    vy = by * x * k * (y * cA / r / rzeta - sA * cA / rzeta - y / r / rz)
    wy = by * x * k * (z * cA / r / rzeta - cA * cA / rzeta + 1.0 / r)
    uz = bz * sA * k * (N * Lzeta - x * x / r / rzeta)
    vz = bz * x * sA * k * (sA / rzeta - y / r / rzeta)
    wz = bz * x * sA * k * (cA / rzeta - z / r / rzeta)
    return ux + uy + uz, vx + vy + vz, wx + wy + wz


def _setup_d(q, alpha, bx, by, bz, nu, vert, side):
    """TDSetupD: into the side frame, AngDisDisp at -pi + alpha, back."""
    s, c = side[1], side[2]
    ry, rz = q[:, 1] - vert[1], q[:, 2] - vert[2]
    y1 = c * ry - s * rz
    z1 = s * ry + c * rz
    by1 = c * by - s * bz
    bz1 = s * by + c * bz
    u, v0, w0 = ang_dis_disp(q[:, 0], y1, z1, -PI + alpha, bx, by1, bz1, nu)
    return u, c * v0 + s * w0, -s * v0 + c * w0


def td_disp_fs(pts, P1, P2, P3, Ss, Ds, Ts, nu):
    """Full-space displacement (n x 3, global axes) at points pts (n x 3)."""
    pts = np.asarray(pts, dtype=np.float64)
    P1, P2, P3 = (np.asarray(P, dtype=np.float64) for P in (P1, P2, P3))
    bx, by, bz = Ts, Ss, Ds
    vn, vs, vd = _frame(P1, P2, P3)
    At = np.array([vn, vs, vd]).T
    q = (pts - P2) @ At
    p1, p2, p3 = (P1 - P2) @ At, np.zeros(3), (P3 - P2) @ At
    e12 = (p2 - p1) / np.linalg.norm(p2 - p1)
    e13 = (p3 - p1) / np.linalg.norm(p3 - p1)
    e23 = (p3 - p2) / np.linalg.norm(p3 - p2)
    A = np.arccos(e12 @ e13)
    B = np.arccos(-e12 @ e23)
    C = np.arccos(e23 @ e13)
    m = _mode(q, p1, p2, p3)
    u = np.zeros((len(q), 3))
    for sg, sel in ((1.0, m == 1), (-1.0, m == -1)):
        if sel.any():
            qs = q[sel]
            parts = [_setup_d(qs, A, bx, by, bz, nu, p1, -sg * e13),
                     _setup_d(qs, B, bx, by, bz, nu, p2, sg * e12),
                     _setup_d(qs, C, bx, by, bz, nu, p3, sg * e23)]
            for j in range(3):
                u[sel, j] = parts[0][j] + parts[1][j] + parts[2][j]
    u[m == 0] = np.nan
    a, b, c = p1 - q, p2 - q, p3 - q
    na, nb, nc = (np.linalg.norm(v, axis=1) for v in (a, b, c))
    Fi = -2.0 * np.arctan2((a * np.cross(b, c)).sum(1),
                           na * nb * nc + (a * b).sum(1) * nc + (a * c).sum(1) * nb
                           + (b * c).sum(1) * na) / (4.0 * PI)
    u = u + np.stack([bx * Fi, by * Fi, bz * Fi], axis=1)
    return u @ np.array([vn, vs, vd])


def td_disp_harmonic(pts, P1, P2, P3, Ss, Ds, Ts, nu):
    """Free-surface correction displacement (TDdisp_HarFunc)."""
    pts = np.asarray(pts, dtype=np.float64)
    P = [np.asarray(x, dtype=np.float64) for x in (P1, P2, P3)]
    vn, vs, vd = _frame(*P)
    Bg = vn * Ts + vs * Ss + vd * Ds
    out = np.zeros((len(pts), 3))
    for PA, PB in ((P[0], P[1]), (P[1], P[2]), (P[2], P[0])):
        side = PB - PA
        beta = np.arccos(-side[2] / np.linalg.norm(side))
        if abs(beta) < 2.2e-10 or abs(PI - beta) < 2.2e-10:
            continue
        e1 = np.array([side[0], side[1], 0.0])
        e1 /= np.linalg.norm(e1)
        e3 = np.array([0.0, 0.0, -1.0])
        e2 = np.cross(e3, e1)
        At = np.array([e1, e2, e3])
        yA = (pts - PA) @ At.T
        yB = yA - side @ At.T
        bv = At @ Bg
        conf = beta * yA[:, 0] >= 0          # configuration I
        v = np.zeros((len(pts), 3))
        for sel, ang in ((conf, -PI + beta), (~conf, beta)):
            if not sel.any():
                continue
            trig = (np.sin(ang), np.cos(ang), 1.0 / np.tan(ang), np.tan(ang / 2.0))
            vA = fsc_displacement(yA[sel, 0], yA[sel, 1], yA[sel, 2], *trig, bv, nu, -PA[2])
            vB = fsc_displacement(yB[sel, 0], yB[sel, 1], yB[sel, 2], *trig, bv, nu, -PB[2])
            for j in range(3):
                v[sel, j] = vB[j] - vA[j]
        out += v @ At
    return out


def td_disp_hs(pts, P1, P2, P3, Ss, Ds, Ts, nu):
    """Half-space displacement (TDdispHS): main + correction + image."""
    P = [np.asarray(x, dtype=np.float64).copy() for x in (P1, P2, P3)]
    um = td_disp_fs(pts, *P, Ss, Ds, Ts, nu)
    uh = td_disp_harmonic(pts, *P, Ss, Ds, Ts, nu)
    Q = [x.copy() for x in P]
    for x in Q:
        x[2] = -x[2]
    ui = td_disp_fs(pts, *Q, Ss, Ds, Ts, nu)
    surface = all(x[2] == 0 for x in Q)
    if surface:
        ui[:, 2] = -ui[:, 2]
    u = um + ui + uh
    return -u if surface else u
