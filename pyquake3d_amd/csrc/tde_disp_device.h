// Displacement of a triangular dislocation (Nikkhoo & Walter 2015), full and
// half space, on the host-computed source records of tde_device.h and
// tde_hs_device.h; the reference's twin is src/DH_greenfunction.py
// (TDdispFS :99-220, TDdispHS :402-443).
#pragma once

#include "tde_device.h"
#include "tde_hs_device.h"

namespace pq3d {

// One angular dislocation at (x, y, z) in its frame, angle with (cA, sA)
// (AngDisDisp, src/DH_greenfunction.py:65-96).
PQ3D_HD inline void ang_dis_disp(double x, double y, double z, double cA, double sA, double bx,
                                 double by, double bz, double nu, double& u, double& v,
                                 double& w) {
  const double eta = y * cA - z * sA;
  double zeta = y * sA + z * cA;
  const double r = sqrt(x * x + y * y + z * z);
  if (zeta > r) zeta = r;  // keeps the logarithms real
  if (z > r) z = r;
  const double k = (1.0 / (8.0 * kPi)) / (1.0 - nu);
  const double N = 1.0 - 2.0 * nu;
  const double rz = r - z, rzeta = r - zeta;
  const double Lz = log(rz), Lzeta = log(rzeta);
  // reciprocals shared by the terms: 1/r, 1/(r - z), 1/(r - zeta)
  const double ir = 1.0 / r, irz = 1.0 / rz, ize = 1.0 / rzeta;
  const double irrz = ir * irz, irze = ir * ize;
  const double ux = bx * k * (x * y * irrz - x * eta * irze);
  const double vx = bx * k * (eta * sA * ize - y * eta * irze + y * y * irrz +
                              N * (cA * Lzeta - Lz));
  const double wx = bx * k * (eta * cA * ize - y * ir - eta * z * irze - N * sA * Lzeta);
  const double uy = by * k * (x * x * cA * irze - x * x * irrz - N * (cA * Lzeta - Lz));
  const double vy = by * x * k * (y * cA * irze - sA * cA * ize - y * irrz);
  const double wy = by * x * k * (z * cA * irze - cA * cA * ize + ir);
  const double uz = bz * sA * k * (N * Lzeta - x * x * irze);
  const double vz = bz * x * sA * k * (sA * ize - y * irze);
  const double wz = bz * x * sA * k * (cA * ize - z * irze);
  u = ux + uy + uz;
  v = vx + vy + vz;
  w = wx + wy + wz;
}

// Full-space displacement (global axes) at one point (TDdispFS for n = 1):
// the three angular dislocations of the point's configuration, NaN on a
// side in the TD plane, plus the Burgers-function term.
PQ3D_HD inline void tde_disp_point(const double* rec, double x, double y, double z, double Ss,
                                   double Ds, double Ts, double nu, double out[3]) {
  double q[3];
  to_td(rec, x, y, z, q);
  const int mode = tde_mode(q[0], q[1], q[2], rec);
  double ut = 0.0, vt = 0.0, wt = 0.0;
  if (mode == 0) {
    ut = vt = wt = __builtin_nan("");
  } else {
    const double sg = (double)mode;
    for (int k = 0; k < 3; ++k) {
      const double s = sg * rec[R_SV + 2 * k], c = sg * rec[R_SV + 2 * k + 1];
      double vy, vz;
      vertex_yz(rec, k, vy, vz);
      const double ry = q[1] - vy, rz = q[2] - vz;
      const double y1 = c * ry - s * rz, z1 = s * ry + c * rz;
      const double by1 = c * Ss - s * Ds, bz1 = s * Ss + c * Ds;
      double u0, v0, w0;
      ang_dis_disp(q[0], y1, z1, rec[R_CS + 2 * k], rec[R_CS + 2 * k + 1], Ts, by1, bz1, nu, u0,
                   v0, w0);
      ut += u0;
      vt += c * v0 + s * w0;
      wt += -s * v0 + c * w0;
    }
  }
  // Burgers function: the solid angle of the triangle seen from the point
  const double a[3] = {-q[0], rec[R_P1] - q[1], rec[R_P1 + 1] - q[2]};
  const double b[3] = {-q[0], -q[1], -q[2]};
  const double cc[3] = {-q[0], rec[R_P3] - q[1], rec[R_P3 + 1] - q[2]};
  const double na = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  const double nb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  const double nc = sqrt(cc[0] * cc[0] + cc[1] * cc[1] + cc[2] * cc[2]);
  const double bxc0 = b[1] * cc[2] - b[2] * cc[1], bxc1 = b[2] * cc[0] - b[0] * cc[2],
               bxc2 = b[0] * cc[1] - b[1] * cc[0];
  const double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  const double ac = a[0] * cc[0] + a[1] * cc[1] + a[2] * cc[2];
  const double bc = b[0] * cc[0] + b[1] * cc[1] + b[2] * cc[2];
  const double Fi = -2.0 * atan2(a[0] * bxc0 + a[1] * bxc1 + a[2] * bxc2,
                                 na * nb * nc + ab * nc + ac * nb + bc * na) / (4.0 * kPi);
  ut += Ts * Fi;
  vt += Ss * Fi;
  wt += Ds * Fi;
  for (int j = 0; j < 3; ++j)
    out[j] = ut * rec[R_VN + j] + vt * rec[R_VS + j] + wt * rec[R_VD + j];
}

// Free-surface correction displacement of one angular dislocation at
// (y1, y2, y3) in its frame, angle with (sinB, cosB, cotB, tan(B/2)), depth
// a of its vertex (AngDisDispFSC, src/DH_greenfunction.py:313-399).
PQ3D_HD inline __attribute__((always_inline)) void fsc_disp(double y1, double y2, double y3,
                                                       const double* trig, double b1, double b2,
                                                       double b3, double nu, double a,
                                                       double v[3]) {
  const double sinB = trig[0], cosB = trig[1], cotB = trig[2], tanH = trig[3];
  const double N1 = 1.0 - 2.0 * nu, M1 = 1.0 - nu;
  const double y3b = y3 + 2.0 * a;
  const double z1b = y1 * cosB + y3b * sinB;
  const double z3b = -y1 * sinB + y3b * cosB;
  const double r2b = y1 * y1 + y2 * y2 + y3b * y3b;
  const double rb = sqrt(r2b);
  const double W6 = rb + y3b, W7 = rb + z3b;
  const double Fib = 2.0 * atan(-y2 / (y1 - W6 / tanH));
  const double L6 = log(W6), L7 = log(W7);
  // reciprocals shared by the terms
  const double irb = 1.0 / rb, iW6 = 1.0 / W6, iW7 = 1.0 / W7, icB = 1.0 / cosB;
  const double ir2b = irb * irb, ir3b = ir2b * irb;
  const double ar = a * irb;
  const double h = y3b - a;
  const double q = rb * cosB + y3b;
  const double c2 = cotB * cotB;
  const double hrb6 = h * irb * iW6, hrb7 = h * irb * iW7;
  const double S = rb * sinB - y1;
  const double u1 = -2.0 * M1 * N1 * Fib * c2 + N1 * y2 * iW6 * ((N1 - ar) * cotB - y1 * iW6 * (nu + ar)) +
                    N1 * y2 * cosB * cotB * iW7 * (cosB + ar) + a * y2 * h * cotB * ir3b +
                    y2 * hrb6 * (-N1 * cotB + y1 * iW6 * (2.0 * nu + ar) + a * y1 * ir2b) +
                    y2 * hrb7 *
                        (cosB * iW7 * (q * (N1 * cosB - ar) * cotB + 2.0 * M1 * S * cosB) -
                         a * y3b * cosB * cotB * ir2b);
  const double u2 = N1 * ((2.0 * M1 * c2 - nu) * L6 - (2.0 * M1 * c2 + N1) * cosB * L7) -
                    N1 * iW6 * (y1 * cotB * (N1 - ar) + nu * y3b - a + y2 * y2 * iW6 * (nu + ar)) -
                    N1 * z1b * cotB * iW7 * (cosB + ar) - a * y1 * h * cotB * ir3b +
                    h * iW6 * (-2.0 * nu + irb * (N1 * y1 * cotB - a) +
                               y2 * y2 * irb * iW6 * (2.0 * nu + ar) + a * y2 * y2 * ir3b) +
                    h * iW7 * (cosB * cosB - irb * (N1 * z1b * cotB + a * cosB) +
                               a * y3b * z1b * cotB * ir3b -
                               irb * iW7 * (y2 * y2 * cosB * cosB - a * z1b * cotB * irb * q));
  const double u3 = 2.0 * M1 * (N1 * Fib * cotB + y2 * iW6 * (2.0 * nu + ar) - y2 * cosB * iW7 * (cosB + ar)) +
                    y2 * h * irb * (2.0 * nu * iW6 + a * ir2b) +
                    y2 * cosB * hrb7 * (N1 - q * iW7 * (cosB + ar) - a * y3b * ir2b);
  const double w1 = N1 * ((2.0 * M1 * c2 + nu) * L6 - (2.0 * M1 * c2 + 1.0) * cosB * L7) +
                    N1 * iW6 * (-N1 * y1 * cotB + nu * y3b - a + a * y1 * cotB * irb + y1 * y1 * iW6 * (nu + ar)) -
                    N1 * cotB * iW7 * (z1b * cosB - a * S * irb * icB) -
                    a * y1 * h * cotB * ir3b +
                    h * iW6 * (2.0 * nu + irb * (N1 * y1 * cotB + a) -
                               y1 * y1 * irb * iW6 * (2.0 * nu + ar) - a * y1 * y1 * ir3b) +
                    h * cotB * iW7 *
                        (-cosB * sinB + a * y1 * y3b * ir3b * icB +
                         S * irb * (2.0 * M1 * cosB - q * iW7 * (1.0 + a * irb * icB)));
  const double w2 = 2.0 * M1 * N1 * Fib * c2 + N1 * y2 * iW6 * (-(N1 - ar) * cotB + y1 * iW6 * (nu + ar)) -
                    N1 * y2 * cotB * iW7 * (1.0 + a * irb * icB) - a * y2 * h * cotB * ir3b +
                    y2 * hrb6 * (N1 * cotB - 2.0 * nu * y1 * iW6 - a * y1 * irb * (irb + iW6)) +
                    y2 * cotB * hrb7 *
                        (-2.0 * M1 * cosB + q * iW7 * (1.0 + a * irb * icB) + a * y3b * ir2b * icB);
  const double w3 = -2.0 * M1 * N1 * cotB * (L6 - cosB * L7) - 2.0 * M1 * y1 * iW6 * (2.0 * nu + ar) +
                    2.0 * M1 * z1b * iW7 * (cosB + ar) +
                    h * irb * (N1 * cotB - 2.0 * nu * y1 * iW6 - a * y1 * ir2b) -
                    h * iW7 * (cosB * sinB + q * cotB * irb * (2.0 * M1 * cosB - q * iW7) +
                               ar * (sinB - y3b * z1b * ir2b - z1b * q * irb * iW7));
  const double s1 = N1 * (y2 * iW6 * (1.0 + ar) - y2 * cosB * iW7 * (cosB + ar)) -
                    y2 * h * irb * (a * ir2b + iW6) +
                    y2 * cosB * hrb7 * (q * iW7 * (cosB + ar) + a * y3b * ir2b);
  const double s2 = N1 * (-sinB * L7 - y1 * iW6 * (1.0 + ar) + z1b * iW7 * (cosB + ar)) +
                    y1 * h * irb * (a * ir2b + iW6) -
                    h * iW7 * (sinB * (cosB - ar) + z1b * irb * (1.0 + a * y3b * ir2b) -
                               irb * iW7 * (y2 * y2 * cosB * sinB - a * z1b * irb * q));
  const double s3 = 2.0 * M1 * Fib + 2.0 * M1 * (y2 * sinB * iW7 * (cosB + ar)) +
                    y2 * sinB * hrb7 * (1.0 + q * iW7 * (cosB + ar) + a * y3b * ir2b);
  const double c = (1.0 / (4.0 * kPi)) / (1.0 - nu);
  v[0] = c * (b1 * u1 + b2 * w1 + b3 * s1);
  v[1] = c * (b1 * u2 + b2 * w2 + b3 * s2);
  v[2] = c * (b1 * u3 + b2 * w3 + b3 * s3);
}

// Correction displacement of the three sides (TDdisp_HarFunc / AngSetupFSC,
// src/DH_greenfunction.py:223-310). A vertical side (beta within 2.2e-10 of
// 0 or pi) contributes nothing, as in the published method.
PQ3D_HD inline void harm_disp_point(const double* h, double x, double y, double z, double Ss,
                                    double Ds, double Ts, double nu, double out[3]) {
  const double* vn = h + H_FRAME;
  const double* vs = h + H_FRAME + 3;
  const double* vd = h + H_FRAME + 6;
  double Bg[3];
  for (int k = 0; k < 3; ++k) Bg[k] = vn[k] * Ts + vs[k] * Ss + vd[k] * Ds;
  out[0] = out[1] = out[2] = 0.0;
  for (int sd = 0; sd < 3; ++sd) {
    const double* s = h + H_SIDE + H_SW * sd;
    if (s[S_VERT] != 0.0 || fabs(s[S_T0]) < 2.2e-10) continue;
    const double ex = s[S_E1], ey = s[S_E1 + 1];
    const double dx = x - s[S_PA], dy = y - s[S_PA + 1], dz = z - s[S_PA + 2];
    const double yA[3] = {dx * ex + dy * ey, dx * ey - dy * ex, -dz};
    const double yB[3] = {yA[0] - s[S_YAB], yA[1] - s[S_YAB + 1], yA[2] - s[S_YAB + 2]};
    const double b1 = Bg[0] * ex + Bg[1] * ey, b2 = Bg[0] * ey - Bg[1] * ex, b3 = -Bg[2];
    // configuration I (yA_x >= 0): angle -pi + beta; otherwise beta
    double trig[4];
    if (yA[0] >= 0.0) {
      trig[0] = -s[S_T0];
      trig[1] = -s[S_T0 + 1];
      trig[2] = s[S_T0 + 2];
      trig[3] = -1.0 / s[S_T0 + 3];
    } else {
      for (int k = 0; k < 4; ++k) trig[k] = s[S_T0 + k];
    }
    double vA[3], vB[3];
    fsc_disp(yA[0], yA[1], yA[2], trig, b1, b2, b3, nu, s[S_AA], vA);
    fsc_disp(yB[0], yB[1], yB[2], trig, b1, b2, b3, nu, s[S_AB], vB);
    const double v0 = vB[0] - vA[0], v1 = vB[1] - vA[1], v2 = vB[2] - vA[2];
    // side frame ey1 = (ex, ey, 0), ey2 = (ey, -ex, 0), ey3 = (0, 0, -1)
    out[0] += v0 * ex + v1 * ey;
    out[1] += v0 * ey - v1 * ex;
    out[2] += -v2;
  }
}

// Half-space displacement (TDdispHS): main + correction + image; for a
// triangle with all vertices at z = 0 the image's vertical component and
// then the whole sum change sign.
PQ3D_HD inline void tde_disp_point_hs(const double* rec, const double* h, double x, double y,
                                      double z, double Ss, double Ds, double Ts, double nu,
                                      double out[3]) {
  double um[3], ui[3], uh[3];
  tde_disp_point(rec, x, y, z, Ss, Ds, Ts, nu, um);
  tde_disp_point(h + H_IMG, x, y, z, Ss, Ds, Ts, nu, ui);
  harm_disp_point(h, x, y, z, Ss, Ds, Ts, nu, uh);
  const bool surf = h[H_SURF] != 0.0;
  if (surf) ui[2] = -ui[2];
  for (int j = 0; j < 3; ++j) {
    const double u = um[j] + ui[j] + uh[j];
    out[j] = surf ? -u : u;
  }
}

}  // namespace pq3d
