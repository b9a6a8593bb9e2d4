// Half-space triangular-dislocation stress (free surface z = 0), FP64,
// host+device.
//
// S = S_main + S_harmonic + S_image (src/TDstressFS_C.cpp:1911-1999):
//   * S_main, S_image: the full-space kernel (tde_device.h) for the triangle
//     and for its mirror image in z = 0; the image's xz and yz components
//     change sign when all vertices lie at z = 0 (:1954-1963);
//   * S_harmonic: free-surface correction of the three angular dislocations
//     along the sides (TDstress_HarFunc :1766-1908, AngSetupFSC_S
//     :1489-1763).
//
// The correction strain is evaluated here as the symmetric gradient of the
// correction displacement of one angular dislocation (Nikkhoo & Walter 2015;
// the reference's displacement twin src/DH_greenfunction.py:313-399), by
// forward-mode differentiation: every quantity carries its value and its
// gradient in the angular dislocation's frame (D3). That is a different
// evaluation order from the reference's expanded strain formula (:613-1486);
// they agree to rounding.
//
// Vertical sides: the reference leaves its outputs unwritten for a side with
// beta within 2.2204e-16 of 0 or pi (:1524-1531), so TDstress_HarFunc adds the
// previous side's stress and strain once more (nothing for the first side).
// harm_* below reproduce that.
#pragma once

#include "tde_device.h"

namespace pq3d {

constexpr int kHsRec = 104;  // doubles per half-space source record

// Half-space record layout (doubles):
//   0..31   full-space record of the image triangle (vertex z mirrored)
//   32      1 if all three vertices lie at z = 0, else 0
//   33..41  side-frame normal, strike, dip (Vn, Vs, Vd of TDstress_HarFunc)
//   42 + 20 e (e = 0, 1, 2: sides P1P2, P2P3, P3P1):
//     +0 vertical flag   +1..3 PA   +4 -PA_z   +5 -PB_z   +6,7 ey1 (x, y)
//     +8..10 PB - PA in the side frame   +11..14 sin, cos, cot, tan(./2) of
//     beta   +15..18 the same of pi - beta
enum : int { H_IMG = 0, H_SURF = 32, H_FRAME = 33, H_SIDE = 42, H_SW = 20 };
enum : int { S_VERT = 0, S_PA = 1, S_AA = 4, S_AB = 5, S_E1 = 6, S_YAB = 8, S_T0 = 11,
             S_T1 = 15 };

// ------------------------------------------------------------ host setup
// Per-triangle half-space record (src/TDstressFS_C.cpp:1766-1816 frame,
// :1505-1582 per side), host side with the C library's trig.
inline void tde_hs_record_host(const double* P1, const double* P2, const double* P3,
                               double* h) {
  for (int k = 0; k < kHsRec; ++k) h[k] = 0.0;
  const double Q1[3] = {P1[0], P1[1], -P1[2]};
  const double Q2[3] = {P2[0], P2[1], -P2[2]};
  const double Q3[3] = {P3[0], P3[1], -P3[2]};
  tde_source_record_host(Q1, Q2, Q3, h + H_IMG);
  h[H_SURF] = (P1[2] == 0.0 && P2[2] == 0.0 && P3[2] == 0.0) ? 1.0 : 0.0;
  const double u[3] = {P2[0] - P1[0], P2[1] - P1[1], P2[2] - P1[2]};
  const double v[3] = {P3[0] - P1[0], P3[1] - P1[1], P3[2] - P1[2]};
  double vn[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2],
                  u[0] * v[1] - u[1] * v[0]};
  const double ln = std::sqrt(vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2]);
  if (ln != 0.0)
    for (double& c : vn) c /= ln;
  double vs[3] = {-vn[1], vn[0], 0.0};  // ez x vn
  const double ls = std::sqrt(vs[0] * vs[0] + vs[1] * vs[1] + vs[2] * vs[2]);
  if (ls == 0.0) {
    vs[0] = 0.0;
    vs[1] = vn[2];
    vs[2] = 0.0;
  } else {
    for (double& c : vs) c /= ls;
  }
  const double vd[3] = {vn[1] * vs[2] - vn[2] * vs[1], vn[2] * vs[0] - vn[0] * vs[2],
                        vn[0] * vs[1] - vn[1] * vs[0]};
  for (int k = 0; k < 3; ++k) {
    h[H_FRAME + k] = vn[k];
    h[H_FRAME + 3 + k] = vs[k];
    h[H_FRAME + 6 + k] = vd[k];
  }
  const double* PA[3] = {P1, P2, P3};
  const double* PB[3] = {P2, P3, P1};
  for (int e = 0; e < 3; ++e) {
    double* s = h + H_SIDE + H_SW * e;
    const double sv[3] = {PB[e][0] - PA[e][0], PB[e][1] - PA[e][1], PB[e][2] - PA[e][2]};
    const double nrm = std::sqrt(sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2]);
    const double beta = std::acos(-sv[2] / nrm);
    const double eps = 2.2204e-16;
    for (int k = 0; k < 3; ++k) s[S_PA + k] = PA[e][k];
    s[S_AA] = -PA[e][2];
    s[S_AB] = -PB[e][2];
    if (std::fabs(beta) < eps || std::fabs(kPi - beta) < eps) {
      s[S_VERT] = 1.0;
      continue;
    }
    double ex = sv[0], ey = sv[1];
    const double l1 = std::sqrt(ex * ex + ey * ey);
    if (l1 != 0.0) {
      ex /= l1;
      ey /= l1;
    }
    s[S_E1] = ex;
    s[S_E1 + 1] = ey;
    // side frame columns ey1 = (ex, ey, 0), ey2 = ey3 x ey1 = (ey, -ex, 0), ey3 = (0, 0, -1)
    s[S_YAB + 0] = sv[0] * ex + sv[1] * ey;
    s[S_YAB + 1] = sv[0] * ey - sv[1] * ex;
    s[S_YAB + 2] = -sv[2];
    const double b1 = kPi - beta;
    s[S_T0 + 0] = std::sin(beta);
    s[S_T0 + 1] = std::cos(beta);
    s[S_T0 + 2] = 1.0 / std::tan(beta);
    s[S_T0 + 3] = std::tan(beta / 2.0);
    s[S_T1 + 0] = std::sin(b1);
    s[S_T1 + 1] = std::cos(b1);
    s[S_T1 + 2] = 1.0 / std::tan(b1);
    s[S_T1 + 3] = std::tan(b1 / 2.0);
  }
}

// ------------------------------------------------- forward-mode numbers
// Value and gradient with respect to (y1, y2, y3).
struct D3 {
  double v, g0, g1, g2;
};
PQ3D_HD inline D3 operator+(D3 a, D3 b) { return {a.v + b.v, a.g0 + b.g0, a.g1 + b.g1, a.g2 + b.g2}; }
PQ3D_HD inline D3 operator-(D3 a, D3 b) { return {a.v - b.v, a.g0 - b.g0, a.g1 - b.g1, a.g2 - b.g2}; }
PQ3D_HD inline D3 operator-(D3 a) { return {-a.v, -a.g0, -a.g1, -a.g2}; }
PQ3D_HD inline D3 operator+(D3 a, double c) { return {a.v + c, a.g0, a.g1, a.g2}; }
PQ3D_HD inline D3 operator+(double c, D3 a) { return {a.v + c, a.g0, a.g1, a.g2}; }
PQ3D_HD inline D3 operator-(D3 a, double c) { return {a.v - c, a.g0, a.g1, a.g2}; }
PQ3D_HD inline D3 operator-(double c, D3 a) { return {c - a.v, -a.g0, -a.g1, -a.g2}; }
PQ3D_HD inline D3 operator*(D3 a, double c) { return {a.v * c, a.g0 * c, a.g1 * c, a.g2 * c}; }
PQ3D_HD inline D3 operator*(double c, D3 a) { return {a.v * c, a.g0 * c, a.g1 * c, a.g2 * c}; }
PQ3D_HD inline D3 operator*(D3 a, D3 b) {
  return {a.v * b.v, a.v * b.g0 + a.g0 * b.v, a.v * b.g1 + a.g1 * b.v, a.v * b.g2 + a.g2 * b.v};
}
PQ3D_HD inline D3 operator/(D3 a, D3 b) {
  const double q = a.v / b.v;
  return {q, (a.g0 - q * b.g0) / b.v, (a.g1 - q * b.g1) / b.v, (a.g2 - q * b.g2) / b.v};
}
PQ3D_HD inline D3 operator/(D3 a, double c) { return {a.v / c, a.g0 / c, a.g1 / c, a.g2 / c}; }
PQ3D_HD inline D3 operator/(double c, D3 b) {
  const double q = c / b.v;
  return {q, -q * b.g0 / b.v, -q * b.g1 / b.v, -q * b.g2 / b.v};
}
PQ3D_HD inline D3 d3sqrt(D3 a) {
  const double r = sqrt(a.v);
  const double k = 0.5 / r;
  return {r, a.g0 * k, a.g1 * k, a.g2 * k};
}
// Terms that enter the displacement only with constant coefficients need
// their gradient alone: log(a) and 2 atan(n/d) carry value 0 here.
// (ia = 1 / a.v, computed once by the caller)
PQ3D_HD inline D3 d3log_grad_r(D3 a, double ia) { return {0.0, a.g0 * ia, a.g1 * ia, a.g2 * ia}; }
// 1 / b
PQ3D_HD inline D3 d3inv(D3 b) {
  const double r = 1.0 / b.v;
  const double m = -r * r;
  return {r, b.g0 * m, b.g1 * m, b.g2 * m};
}
PQ3D_HD inline D3 d3atan2x2_grad(D3 n, D3 d) {
  const double w = 2.0 / (n.v * n.v + d.v * d.v);
  return {0.0, (d.v * n.g0 - n.v * d.g0) * w, (d.v * n.g1 - n.v * d.g1) * w,
          (d.v * n.g2 - n.v * d.g2) * w};
}

// Gradient of the correction displacement of one angular dislocation at y
// (its frame) for the unit Burgers component e_K: g[i][j] = d u_i / d y_j,
// including the 1/(4 pi (1 - nu)) factor. Formulas: Nikkhoo & Walter (2015)
// free-surface correction displacement, as in
// src/DH_greenfunction.py:313-399. One component per call, kept out of line,
// so that the forward-mode temporaries of one call fit the register file.
template <int K>
PQ3D_HD inline __attribute__((noinline)) void fsc_grad(double y1v, double y2v, double y3v,
                                               const double* t, double nu, double a,
                                               double g[3][3]) {
  const double sinB = t[0], cosB = t[1], cotB = t[2], tanH = t[3];
  const D3 y1 = {y1v, 1.0, 0.0, 0.0};
  const D3 y2 = {y2v, 0.0, 1.0, 0.0};
  const D3 y3 = {y3v, 0.0, 0.0, 1.0};
  const double N1 = 1.0 - 2.0 * nu, M1 = 1.0 - nu;
  const D3 y3b = y3 + 2.0 * a;
  const D3 z1b = y1 * cosB + y3b * sinB;
  const D3 z3b = (-y1) * sinB + y3b * cosB;
  const D3 r2b = y1 * y1 + y2 * y2 + y3b * y3b;
  const D3 rb = d3sqrt(r2b);
  const D3 W6 = rb + y3b;
  const D3 W7 = rb + z3b;
  const D3 hh = y3b - a;
  const D3 q = rb * cosB + y3b;
  const double c2 = cotB * cotB;
  // reciprocals shared by the terms (one division each instead of four per
  // forward-mode quotient)
  const D3 irb = d3inv(rb), iW6 = d3inv(W6), iW7 = d3inv(W7);
  const D3 ir2b = irb * irb, ir3b = ir2b * irb;
  const D3 irbW6 = irb * iW6, irbW7 = irb * iW7;
  const D3 ar = a * irb;
  const double icB = 1.0 / cosB, itH = 1.0 / tanH;
  const D3 S = rb * sinB - y1;
  D3 o1, o2, o3;
  if constexpr (K == 0) {
    const D3 Fib = d3atan2x2_grad(-y2, (-W6) * itH + y1);
    const D3 L6 = d3log_grad_r(W6, iW6.v), L7 = d3log_grad_r(W7, iW7.v);
    o1 = (-2.0 * M1 * N1 * c2) * Fib + N1 * y2 * iW6 * ((N1 - ar) * cotB - y1 * iW6 * (nu + ar)) +
         (N1 * cosB * cotB) * y2 * iW7 * (cosB + ar) + (a * cotB) * y2 * hh * ir3b +
         y2 * hh * irbW6 * ((-N1 * cotB) + y1 * iW6 * (2.0 * nu + ar) + a * y1 * ir2b) +
         y2 * hh * irbW7 *
             (cosB * iW7 * (q * (N1 * cosB - ar) * cotB + (2.0 * M1 * cosB) * S) -
              (a * cosB * cotB) * y3b * ir2b);
    const D3 y2s = y2 * y2;
    o2 = N1 * ((2.0 * M1 * c2 - nu) * L6 - ((2.0 * M1 * c2 + N1) * cosB) * L7) -
         N1 * iW6 * (cotB * y1 * (N1 - ar) + nu * y3b - a + y2s * iW6 * (nu + ar)) -
         (N1 * cotB) * z1b * iW7 * (cosB + ar) - (a * cotB) * y1 * hh * ir3b +
         hh * iW6 *
             ((-2.0 * nu) + irb * (N1 * cotB * y1 - a) + y2s * irbW6 * (2.0 * nu + ar) +
              a * y2s * ir3b) +
         hh * iW7 *
             (cosB * cosB - irb * (N1 * cotB * z1b + a * cosB) +
              (a * cotB) * y3b * z1b * ir3b -
              irbW7 * ((cosB * cosB) * y2s - (a * cotB) * z1b * irb * q));
    o3 = 2.0 * M1 * ((N1 * cotB) * Fib + y2 * iW6 * (2.0 * nu + ar) - cosB * y2 * iW7 * (cosB + ar)) +
         y2 * hh * irb * ((2.0 * nu) * iW6 + a * ir2b) +
         cosB * y2 * hh * irbW7 * (N1 - q * iW7 * (cosB + ar) - a * y3b * ir2b);
  } else if constexpr (K == 1) {
    const D3 Fib = d3atan2x2_grad(-y2, (-W6) * itH + y1);
    const D3 L6 = d3log_grad_r(W6, iW6.v), L7 = d3log_grad_r(W7, iW7.v);
    const D3 y1s = y1 * y1;
    o1 = N1 * ((2.0 * M1 * c2 + nu) * L6 - ((2.0 * M1 * c2 + 1.0) * cosB) * L7) +
         N1 * iW6 * ((-N1 * cotB) * y1 + nu * y3b - a + (a * cotB) * y1 * irb + y1s * iW6 * (nu + ar)) -
         (N1 * cotB) * iW7 * (cosB * z1b - a * S * irb * icB) -
         (a * cotB) * y1 * hh * ir3b +
         hh * iW6 *
             ((2.0 * nu) + irb * (N1 * cotB * y1 + a) - y1s * irbW6 * (2.0 * nu + ar) -
              a * y1s * ir3b) +
         cotB * hh * iW7 *
             ((-cosB * sinB) + a * y1 * y3b * ir3b * icB +
              S * irb * ((2.0 * M1 * cosB) - q * iW7 * (1.0 + a * irb * icB)));
    o2 = (2.0 * M1 * N1 * c2) * Fib + N1 * y2 * iW6 * ((-(N1 - ar)) * cotB + y1 * iW6 * (nu + ar)) -
         (N1 * cotB) * y2 * iW7 * (1.0 + a * irb * icB) - (a * cotB) * y2 * hh * ir3b +
         y2 * hh * irbW6 * (N1 * cotB - (2.0 * nu) * y1 * iW6 - a * y1 * irb * (irb + iW6)) +
         cotB * y2 * hh * irbW7 *
             ((-2.0 * M1 * cosB) + q * iW7 * (1.0 + a * irb * icB) + a * y3b * ir2b * icB);
    o3 = (-2.0 * M1 * N1 * cotB) * (L6 - cosB * L7) - (2.0 * M1) * y1 * iW6 * (2.0 * nu + ar) +
         (2.0 * M1) * z1b * iW7 * (cosB + ar) +
         hh * irb * (N1 * cotB - (2.0 * nu) * y1 * iW6 - a * y1 * ir2b) -
         hh * iW7 *
             (cosB * sinB + cotB * q * irb * ((2.0 * M1 * cosB) - q * iW7) +
              ar * (sinB - y3b * z1b * ir2b - z1b * q * irbW7));
  } else {
    const D3 Fib = d3atan2x2_grad(-y2, (-W6) * itH + y1);
    const D3 L7 = d3log_grad_r(W7, iW7.v);
    o1 = N1 * (y2 * iW6 * (1.0 + ar) - cosB * y2 * iW7 * (cosB + ar)) -
         y2 * hh * irb * (a * ir2b + iW6) +
         cosB * y2 * hh * irbW7 * (q * iW7 * (cosB + ar) + a * y3b * ir2b);
    o2 = N1 * ((-sinB) * L7 - y1 * iW6 * (1.0 + ar) + z1b * iW7 * (cosB + ar)) +
         y1 * hh * irb * (a * ir2b + iW6) -
         hh * iW7 *
             (sinB * (cosB - ar) + z1b * irb * (1.0 + a * y3b * ir2b) -
              irbW7 * ((cosB * sinB) * (y2 * y2) - a * z1b * irb * q));
    o3 = (2.0 * M1) * Fib + (2.0 * M1 * sinB) * y2 * iW7 * (cosB + ar) +
         sinB * y2 * hh * irbW7 * (1.0 + q * iW7 * (cosB + ar) + a * y3b * ir2b);
  }
  const double c = 1.0 / 4.0 / kPi / (1.0 - nu);
  g[0][0] = c * o1.g0;
  g[0][1] = c * o1.g1;
  g[0][2] = c * o1.g2;
  g[1][0] = c * o2.g0;
  g[1][1] = c * o2.g1;
  g[1][2] = c * o2.g2;
  g[2][0] = c * o3.g0;
  g[2][1] = c * o3.g1;
  g[2][2] = c * o3.g2;
}

// Side frame -> global: T' = A T A^T, A columns (ey1, ey2, ey3) =
// ((ex, ey, 0), (ey, -ex, 0), (0, 0, -1)).
PQ3D_HD inline void side_to_global(double ex, double ey, const double T[6], double o[6]) {
  const double A[3][3] = {{ex, ey, 0.0}, {ey, -ex, 0.0}, {0.0, 0.0, -1.0}};
  const double S[3][3] = {{T[0], T[3], T[4]}, {T[3], T[1], T[5]}, {T[4], T[5], T[2]}};
  double R[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) acc += A[i][k] * S[k][l] * A[j][l];
      R[i][j] = acc;
    }
  o[0] = R[0][0];
  o[1] = R[1][1];
  o[2] = R[2][2];
  o[3] = R[0][1];
  o[4] = R[0][2];
  o[5] = R[1][2];
}

// Accumulate bv_r[K] (g_B - g_A) for the unit component K into M_r.
template <int K, int NB>
PQ3D_HD inline void harm_component(const double yA[3], const double yB[3], double f,
                                   const double* trig, double nu, double aA, double aB,
                                   const double (&bv)[NB][3], double (&M)[NB][3][3]) {
  double gA[3][3], gB[3][3];
  fsc_grad<K>(f * yA[0], f * yA[1], yA[2], trig, nu, aA, gA);
  fsc_grad<K>(f * yB[0], f * yB[1], yB[2], trig, nu, aB, gB);
  for (int r = 0; r < NB; ++r)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) M[r][i][j] += bv[r][K] * (gB[i][j] - gA[i][j]);
}

// Correction strain (global) of side s at point (x, y, z) for NB global
// Burgers vectors Bg[NB][3] (AngSetupFSC_S, src/TDstressFS_C.cpp:1533-1735).
template <int NB>
PQ3D_HD inline void harm_side(const double* s, double x, double y, double z,
                              const double (&Bg)[NB][3], double nu, double (&e)[NB][6]) {
  const double ex = s[S_E1], ey = s[S_E1 + 1];
  const double dx = x - s[S_PA], dy = y - s[S_PA + 1], dz = z - s[S_PA + 2];
  const double yA[3] = {dx * ex + dy * ey, dx * ey - dy * ex, -dz};
  const double yB[3] = {yA[0] - s[S_YAB], yA[1] - s[S_YAB + 1], yA[2] - s[S_YAB + 2]};
  // configuration I (beta * yA_x >= 0, beta > 0): mirrored y1, y2 and b1, b2,
  // angle pi - beta; the xz and yz strains change sign back
  const bool conf = yA[0] >= 0.0;
  const double f = conf ? -1.0 : 1.0;
  const double* trig = s + (conf ? S_T1 : S_T0);
  double bv[NB][3];
  double M[NB][3][3];
  for (int r = 0; r < NB; ++r) {
    bv[r][0] = f * (Bg[r][0] * ex + Bg[r][1] * ey);
    bv[r][1] = f * (Bg[r][0] * ey - Bg[r][1] * ex);
    bv[r][2] = -Bg[r][2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) M[r][i][j] = 0.0;
  }
  harm_component<0, NB>(yA, yB, f, trig, nu, s[S_AA], s[S_AB], bv, M);
  harm_component<1, NB>(yA, yB, f, trig, nu, s[S_AA], s[S_AB], bv, M);
  harm_component<2, NB>(yA, yB, f, trig, nu, s[S_AA], s[S_AB], bv, M);
  for (int r = 0; r < NB; ++r) {
    const double d[6] = {M[r][0][0],
                         M[r][1][1],
                         M[r][2][2],
                         0.5 * (M[r][0][1] + M[r][1][0]),
                         f * 0.5 * (M[r][0][2] + M[r][2][0]),
                         f * 0.5 * (M[r][1][2] + M[r][2][1])};
    side_to_global(ex, ey, d, e[r]);
  }
}

// Correction stress and strain of the three sides for NB Burgers vectors
// given as (Ss, Ds, Ts) (TDstress_HarFunc), with the reference's reuse of the
// previous side's values for a vertical side.
template <int NB>
PQ3D_HD inline void harm_point(const double* h, double x, double y, double z,
                               const double (&bst)[NB][3], double mu, double lam, double nu,
                               double (&stress)[NB][6], double (&strain)[NB][6]) {
  const double* vn = h + H_FRAME;
  const double* vs = h + H_FRAME + 3;
  const double* vd = h + H_FRAME + 6;
  double Bg[NB][3];
  for (int r = 0; r < NB; ++r)
    for (int k = 0; k < 3; ++k)
      Bg[r][k] = vn[k] * bst[r][2] + vs[k] * bst[r][0] + vd[k] * bst[r][1];
  double prev[NB][6];
  for (int r = 0; r < NB; ++r)
    for (int k = 0; k < 6; ++k) {
      prev[r][k] = 0.0;
      strain[r][k] = 0.0;
    }
  for (int sd = 0; sd < 3; ++sd) {
    const double* s = h + H_SIDE + H_SW * sd;
    if (s[S_VERT] == 0.0) harm_side<NB>(s, x, y, z, Bg, nu, prev);
    for (int r = 0; r < NB; ++r)
      for (int k = 0; k < 6; ++k) strain[r][k] += prev[r][k];
  }
  for (int r = 0; r < NB; ++r) hooke(strain[r], mu, lam, stress[r]);
}

// Full-space point evaluation kept out of line in the half-space paths (the
// harmonic part already fills the register file).
PQ3D_HD inline __attribute__((noinline)) void tde_point_ol(const double* rec, double x, double y,
                                                   double z, double Ss, double Ds, double Ts,
                                                   double mu, double lam, double nu,
                                                   double stress[6], double strain[6]) {
  tde_point(rec, x, y, z, Ss, Ds, Ts, mu, lam, nu, stress, strain);
}

// General Burgers vector: half-space stress and strain at one point.
PQ3D_HD inline void tde_point_hs(const double* rec, const double* h, double x, double y,
                                 double z, double Ss, double Ds, double Ts, double mu, double lam,
                                 double nu, double stress[6], double strain[6]) {
  double sm[6], em[6], si[6], ei[6];
  tde_point_ol(rec, x, y, z, Ss, Ds, Ts, mu, lam, nu, sm, em);
  tde_point_ol(h + H_IMG, x, y, z, Ss, Ds, Ts, mu, lam, nu, si, ei);
  if (h[H_SURF] != 0.0) {
    si[4] = -si[4];
    si[5] = -si[5];
    ei[4] = -ei[4];
    ei[5] = -ei[5];
  }
  const double b[1][3] = {{Ss, Ds, Ts}};
  double sh[1][6], eh[1][6];
  harm_point<1>(h, x, y, z, b, mu, lam, nu, sh, eh);
  for (int k = 0; k < 6; ++k) {
    stress[k] = sm[k] + sh[0][k] + si[k];
    strain[k] = em[k] + eh[0][k] + ei[k];
  }
}

// Fused full-space strike + dip pair kept out of line in the half-space path.
PQ3D_HD inline __attribute__((noinline)) void tde_pair_traction_sd_ol(
    const double* rec, const double* p, const double* f, double mu, double lam, double nu,
    double ts[3], double td[3]) {
  tde_pair_traction_sd(rec, p, f, mu, lam, nu, ts, td);
}

// Unit strike and unit dip slip, half space, projected on the receiver frame
// f: ts = (A1s, A2s, Bs), td = (A1d, A2d, Bd). Tractions are linear in the
// stress, so main, image and correction are projected separately; main and
// image use the fused strike + dip evaluation. A triangle with all vertices
// at z = 0 (its image's xz and yz components change sign) takes the
// component-wise path.
PQ3D_HD inline void tde_pair_traction_sd_hs(const double* rec, const double* h, const double* p,
                                            const double* f, double mu, double lam, double nu,
                                            double ts[3], double td[3]) {
  const double b[2][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}};
  double sh[2][6], eh[2][6];
  if (h[H_SURF] == 0.0) {
    double tsm[3], tdm[3], tsi[3], tdi[3], tsh[3], tdh[3];
    tde_pair_traction_sd_ol(rec, p, f, mu, lam, nu, tsm, tdm);
    tde_pair_traction_sd_ol(h + H_IMG, p, f, mu, lam, nu, tsi, tdi);
    harm_point<2>(h, p[0], p[1], p[2], b, mu, lam, nu, sh, eh);
    project(sh[0], f, tsh);
    project(sh[1], f, tdh);
    for (int k = 0; k < 3; ++k) {
      ts[k] = tsm[k] + tsh[k] + tsi[k];
      td[k] = tdm[k] + tdh[k] + tdi[k];
    }
    return;
  }
  double sm[6], em[6], dm[6], fm[6], si[6], ei[6], di[6], fi[6];
  tde_point_ol(rec, p[0], p[1], p[2], 1.0, 0.0, 0.0, mu, lam, nu, sm, em);
  tde_point_ol(rec, p[0], p[1], p[2], 0.0, 1.0, 0.0, mu, lam, nu, dm, fm);
  tde_point_ol(h + H_IMG, p[0], p[1], p[2], 1.0, 0.0, 0.0, mu, lam, nu, si, ei);
  tde_point_ol(h + H_IMG, p[0], p[1], p[2], 0.0, 1.0, 0.0, mu, lam, nu, di, fi);
  si[4] = -si[4];
  si[5] = -si[5];
  di[4] = -di[4];
  di[5] = -di[5];
  harm_point<2>(h, p[0], p[1], p[2], b, mu, lam, nu, sh, eh);
  double s1[6], s2[6];
  for (int k = 0; k < 6; ++k) {
    s1[k] = sm[k] + sh[0][k] + si[k];
    s2[k] = dm[k] + sh[1][k] + di[k];
  }
  project(s1, f, ts);
  project(s2, f, td);
}

}  // namespace pq3d
