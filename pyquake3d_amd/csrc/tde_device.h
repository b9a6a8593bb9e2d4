// Full-space triangular-dislocation (TD) stress, FP64, host+device.
//
// Math: Nikkhoo & Walter (2015), artefact-free full-space solution, as the
// reference evaluates it (src/TDstressFS_C.cpp:51-521).
//
// Parity design. The reference's results near the singular lines of the
// angular dislocations come from heavy cancellation, so reproducing them to
// 1e-10 requires the same IEEE operation sequence, not just the same formulas:
//   * every expression inside the strain brackets keeps the reference's
//     association and its true divisions (no reciprocal rewriting); the library
//     is compiled with FP contraction off (no FMA), as the x86-64 reference is;
//   * the only transcendental calls (acos, cos, sin) are per triangle: they
//     are evaluated once per source, on the host, with the C library, into a
//     32-double "source record" (tde_source_record); the device only needs
//     sqrt and division, which are correctly rounded on both sides.
//
// Throughput design (MI355X):
//   * per-source work is hoisted out of the (receiver, source) pair loop into
//     the record (the reference redoes it for every call, :292-395);
//   * unit strike and unit dip slip share one pass over the geometry: each
//     strain bracket is computed once and multiplied by the Burgers-dependent
//     prefactor of either slip, which is bitwise what two separate reference
//     evaluations give (tde_pair_traction_sd);
//   * the Ts (bx) brackets are skipped when Ts = 0: they are multiplied by a
//     zero prefactor there, so they only matter for NaN propagation, and every
//     denominator they divide by also divides a by/bz bracket of the same
//     component (so a point that is non-finite in the reference is non-finite
//     here too).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#define PQ3D_HD __host__ __device__

namespace pq3d {

constexpr double kPi = 3.141592653589793;
constexpr int kSrcRec = 32;  // doubles per source record (256 B)

// Source record layout (doubles):
//  0..2  normal (vn)      3..5 strike (vs)     6..8 dip (vd)
//  9..11 P2 (origin of TD coordinates)
// 12,13 p1 (y,z) in TD coordinates        14,15 p3 (y,z)
// 16    unused
// 17,18 side 0 direction (y,z) = -e13     19,20 side 1 = e12    21,22 side 2 = e23
// 23,24 cos, sin of (-pi + angle A)       25,26 same for B      27,28 same for C
// 29..31 unused
enum : int {
  R_VN = 0, R_VS = 3, R_VD = 6, R_P2 = 9, R_P1 = 12, R_P3 = 14,
  R_SV = 17, R_CS = 23
};

// Per-triangle setup (src/TDstressFS_C.cpp:292-395), host side (C library
// trig). A horizontal triangle divides by the zero strike norm exactly as the
// reference does (:317-325) and produces a non-finite record.
inline void tde_source_record_host(const double* P1, const double* P2, const double* P3,
                                   double* rec) {
  double vn[3];
  vn[0] = (P2[1] - P1[1]) * (P3[2] - P1[2]) - (P2[2] - P1[2]) * (P3[1] - P1[1]);
  vn[1] = (P2[2] - P1[2]) * (P3[0] - P1[0]) - (P2[0] - P1[0]) * (P3[2] - P1[2]);
  vn[2] = (P2[0] - P1[0]) * (P3[1] - P1[1]) - (P2[1] - P1[1]) * (P3[0] - P1[0]);
  double nn = std::sqrt(vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2]);
  vn[0] /= nn; vn[1] /= nn; vn[2] /= nn;
  // strike = eZ x normal, dip = normal x strike
  double vs[3] = {0.0 * vn[2] - 1.0 * vn[1], 1.0 * vn[0] - 0.0 * vn[2], 0.0 * vn[1] - 0.0 * vn[0]};
  nn = std::sqrt(vs[0] * vs[0] + vs[1] * vs[1] + vs[2] * vs[2]);
  if (nn == 0) { vs[0] = 0.0 * vn[2]; vs[1] = 1.0 * vn[2]; vs[2] = 0.0 * vn[2]; }
  vs[0] /= nn; vs[1] /= nn; vs[2] /= nn;
  const double vd[3] = {vn[1] * vs[2] - vn[2] * vs[1], vn[2] * vs[0] - vn[0] * vs[2],
                        vn[0] * vs[1] - vn[1] * vs[0]};
  // vertices in TD coordinates (origin P2, axes normal/strike/dip)
  double p1[3], p3[3];
  for (int i = 0; i < 3; ++i) {
    const double a = (i == 0) ? vn[0] : (i == 1) ? vs[0] : vd[0];
    const double b = (i == 0) ? vn[1] : (i == 1) ? vs[1] : vd[1];
    const double c = (i == 0) ? vn[2] : (i == 1) ? vs[2] : vd[2];
    p1[i] = a * (P1[0] - P2[0]) + b * (P1[1] - P2[1]) + c * (P1[2] - P2[2]);
    p3[i] = a * (P3[0] - P2[0]) + b * (P3[1] - P2[1]) + c * (P3[2] - P2[2]);
  }
  double e12[3] = {0.0 - p1[0], 0.0 - p1[1], 0.0 - p1[2]};
  double e13[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
  double e23[3] = {p3[0] - 0.0, p3[1] - 0.0, p3[2] - 0.0};
  double l;
  l = std::sqrt(e12[0] * e12[0] + e12[1] * e12[1] + e12[2] * e12[2]);
  e12[0] /= l; e12[1] /= l; e12[2] /= l;
  l = std::sqrt(e13[0] * e13[0] + e13[1] * e13[1] + e13[2] * e13[2]);
  e13[0] /= l; e13[1] /= l; e13[2] /= l;
  l = std::sqrt(e23[0] * e23[0] + e23[1] * e23[1] + e23[2] * e23[2]);
  e23[0] /= l; e23[1] /= l; e23[2] /= l;
  const double aA = std::acos(e12[0] * e13[0] + e12[1] * e13[1] + e12[2] * e13[2]);
  const double aB = std::acos((-e12[0]) * e23[0] + (-e12[1]) * e23[1] + (-e12[2]) * e23[2]);
  const double aC = std::acos(e23[0] * e13[0] + e23[1] * e13[1] + e23[2] * e13[2]);
  for (int k = 0; k < 3; ++k) {
    rec[R_VN + k] = vn[k]; rec[R_VS + k] = vs[k]; rec[R_VD + k] = vd[k]; rec[R_P2 + k] = P2[k];
  }
  rec[R_P1] = p1[1]; rec[R_P1 + 1] = p1[2];
  rec[R_P3] = p3[1]; rec[R_P3 + 1] = p3[2];
  rec[16] = 0.0;
  rec[R_SV + 0] = -e13[1]; rec[R_SV + 1] = -e13[2];
  rec[R_SV + 2] = e12[1];  rec[R_SV + 3] = e12[2];
  rec[R_SV + 4] = e23[1];  rec[R_SV + 5] = e23[2];
  const double al[3] = {-kPi + aA, -kPi + aB, -kPi + aC};
  for (int k = 0; k < 3; ++k) { rec[R_CS + 2 * k] = std::cos(al[k]); rec[R_CS + 2 * k + 1] = std::sin(al[k]); }
  rec[29] = 0.0; rec[30] = 0.0; rec[31] = 0.0;
}

// Configuration of a TD-coordinate point (q0 normal, q1 strike, q2 dip
// components): +1, -1, or 0 (on a side and in the TD plane), with the exact
// comparisons of src/TDstressFS_C.cpp:125-166 (p2 is the origin).
PQ3D_HD inline int tde_mode(double q0, double q1, double q2, const double* rec) {
  const double x1 = rec[R_P1], y1 = rec[R_P1 + 1];
  const double x2 = 0.0, y2 = 0.0;
  const double x3 = rec[R_P3], y3 = rec[R_P3 + 1];
  const double den = (x1 - x3) * (y2 - y3) - (x2 - x3) * (y1 - y3);
  const double a = ((q1 - x3) * (y2 - y3) - (x2 - x3) * (q2 - y3)) / den;
  const double b = ((x1 - x3) * (q2 - y3) - (q1 - x3) * (y1 - y3)) / den;
  const double c = 1.0 - a - b;
  int mode = 1;
  if ((a <= 0) && (b > c) && (c > a)) mode = -1;
  else if ((b <= 0) && (c > a) && (a > b)) mode = -1;
  else if ((c <= 0) && (a > b) && (b > c)) mode = -1;
  else if ((a == 0) && (b >= 0) && (c >= 0)) mode = 0;
  else if ((a >= 0) && (b == 0) && (c >= 0)) mode = 0;
  else if ((a >= 0) && (b >= 0) && (c == 0)) mode = 0;
  if (mode == 0 && q0 != 0) mode = 1;
  return mode;
}

// Strain brackets of one angular dislocation (src/TDstressFS_C.cpp:68-119).
// For component k the reference strain is
//   E_k = lead_k + sx_k * bx/8/pi/(1-nu) * X_k + sy_k * by*x/8/pi/(1-nu) * Y_k
//                + sz_k * bz*x*sinA/8/pi/(1-nu) * Z_k
// (the xy and xz components use by/..., bz*sinA/... without x); X are only
// needed for Ts != 0.
struct AdBrackets {
  double fx, fy, fz;      // rFi_rx, rFi_ry, rFi_rz
  double Y[6], Z[6];      // by and bz brackets
  double X[6];            // bx brackets (only filled when with_bx)
  double x, sinA;
};

PQ3D_HD inline void ad_brackets(double x, double y, double z, double cosA, double sinA, double nu,
                                bool with_bx, AdBrackets& o) {
  const double eta = y * cosA - z * sinA;
  const double zeta = y * sinA + z * cosA;
  const double x2 = x * x, y2 = y * y, z2 = z * z;
  const double r2 = x2 + y2 + z2;
  const double r = sqrt(r2);
  const double r3 = r2 * r;
  const double rz = r * (r - z);
  const double r2z2 = r2 * (r - z) * (r - z);
  const double r3z = r3 * (r - z);
  const double W = zeta - r;
  const double W2 = W * W;
  const double Wr = W * r;
  const double W2r = W2 * r;
  const double Wr3 = W * r3;
  const double W2r2 = W2 * r2;
  const double C = (r * cosA - z) / Wr;
  const double S = (r * sinA - y) / Wr;
  o.fx = (eta / r / (r - zeta) - y / r / (r - z)) / 4.0 / kPi;
  o.fy = (x / r / (r - z) - cosA * x / r / (r - zeta)) / 4.0 / kPi;
  o.fz = (sinA * x / r / (r - zeta)) / 4.0 / kPi;
  o.x = x;
  o.sinA = sinA;
  // by brackets (xx, yy, zz, yz carry the factor by*x; xy, xz carry by)
  o.Y[0] = ((2 * nu + 1) / Wr + x2 / W2r2 - x2 / Wr3) * cosA + (2 * nu + 1) / rz - x2 / r2z2 - x2 / r3z;
  o.Y[1] = 1.0 / rz - y2 / r2z2 - y2 / r3z + (1.0 / Wr + S * S - y2 / Wr3) * cosA;
  o.Y[2] = (1.0 / Wr + C * C - z2 / Wr3) * cosA + z / r3;
  o.Y[3] = x2 * y / r2z2 - nu * y / rz + x2 * y / r3z + nu * cosA * S + x2 * y * cosA / Wr3 + x2 * cosA * S / Wr;
  o.Y[4] = -x2 / r3 + nu / r + nu * cosA * C + x2 * z * cosA / Wr3 + x2 * cosA * C / Wr;
  o.Y[5] = y / r3 + sinA * cosA * cosA / W2 - cosA * (y * cosA + z * sinA) / W2r + y * z * cosA / W2r2 - y * z * cosA / Wr3;
  // bz brackets (xx, yy, zz, yz carry bz*x*sinA; xy, xz carry bz*sinA)
  o.Z[0] = (2 * nu + 1) / Wr + x2 / W2r2 - x2 / Wr3;
  o.Z[1] = 1.0 / Wr + S * S - y2 / Wr3;
  o.Z[2] = 1.0 / Wr + C * C - z2 / Wr3;
  o.Z[3] = nu * S + x2 * S / Wr + x2 * y / Wr3;
  o.Z[4] = nu * C + x2 * C / Wr + x2 * z / Wr3;
  o.Z[5] = y * z / Wr3 - sinA * cosA / W2 + (y * cosA + z * sinA) / W2r - y * z / W2r2;
  if (with_bx) {
    o.X[0] = eta / Wr + eta * x2 / W2r2 - eta * x2 / Wr3 + y / rz - x2 * y / r2z2 - x2 * y / r3z;
    o.X[1] = (1.0 / Wr + S * S - y2 / Wr3) * eta + (2 * nu + 1) * y / rz - y2 * y / r2z2 - y2 * y / r3z - 2 * nu * cosA * S;
    o.X[2] = eta / W / r + eta * C * C - eta * z2 / Wr3 + y * z / r3 + 2 * nu * sinA * C;
    o.X[3] = x * y2 / r2z2 - nu * x / rz + x * y2 / r3z - nu * x * cosA / Wr + eta * x * S / Wr + eta * x * y / Wr3;
    o.X[4] = -x * y / r3 + nu * x * sinA / Wr + eta * x * C / Wr + eta * x * z / Wr3;
    o.X[5] = y2 / r3 - nu / r - nu * cosA * C + nu * sinA * S + eta * sinA * cosA / W2 - eta * (y * cosA + z * sinA) / W2r + eta * y * z / W2r2 - eta * y * z / Wr3;
  }
}

// Combine brackets into the strain for Burgers vector (bx, by, bz) with the
// reference's evaluation order (src/TDstressFS_C.cpp:90-119).
PQ3D_HD inline void ad_combine(const AdBrackets& o, double bx, double by, double bz, double nu,
                               bool with_bx, double e[6]) {
  const double x = o.x, sinA = o.sinA;
  const double py = by * x / 8.0 / kPi / (1 - nu);          // by*x/8/pi/(1-nu)
  const double pz = bz * x * sinA / 8.0 / kPi / (1 - nu);   // bz*x*sinA/8/pi/(1-nu)
  const double qy = by / 8.0 / kPi / (1 - nu);              // by/8/pi/(1-nu)
  const double qz = bz * sinA / 8.0 / kPi / (1 - nu);       // bz*sinA/8/pi/(1-nu)
  if (with_bx) {
    const double px = bx / 8.0 / kPi / (1 - nu);
    e[0] = bx * o.fx + px * o.X[0] - py * o.Y[0] + pz * o.Z[0];
    e[1] = by * o.fy + px * o.X[1] - py * o.Y[1] + pz * o.Z[1];
    e[2] = bz * o.fz + px * o.X[2] - py * o.Y[2] + pz * o.Z[2];
    e[3] = bx * o.fy / 2 + by * o.fx / 2 - px * o.X[3] + qy * o.Y[3] - qz * o.Z[3];
    e[4] = bx * o.fz / 2 + bz * o.fx / 2 - px * o.X[4] + qy * o.Y[4] - qz * o.Z[4];
    e[5] = by * o.fz / 2 + bz * o.fy / 2 + px * o.X[5] - py * o.Y[5] - pz * o.Z[5];
  } else {
    // bx = 0: the reference's leading zero terms drop out exactly
    e[0] = -(py * o.Y[0]) + pz * o.Z[0];
    e[1] = by * o.fy - py * o.Y[1] + pz * o.Z[1];
    e[2] = bz * o.fz - py * o.Y[2] + pz * o.Z[2];
    e[3] = by * o.fx / 2 + qy * o.Y[3] - qz * o.Z[3];
    e[4] = bz * o.fx / 2 + qy * o.Y[4] - qz * o.Z[4];
    e[5] = by * o.fz / 2 + bz * o.fy / 2 - py * o.Y[5] - pz * o.Z[5];
  }
}

// Rotate a side-frame strain back to TD coordinates and accumulate:
// T' = B^T T B with B = [[1,0,0],[0,c,-s],[0,s,c]], written with the
// reference's term order (src/TDstressFS_C.cpp:21-49 with the zero entries of
// B dropped; they only multiply finite values there or feed NaN into a point
// that is already non-finite).
PQ3D_HD inline void side_rot(const double T[6], double c, double s, double o[6]) {
  const double xx = T[0], yy = T[1], zz = T[2], xy = T[3], xz = T[4], yz = T[5];
  o[0] = xx;
  o[1] = 2 * c * s * yz + c * c * yy + s * s * zz;
  o[2] = 2 * (-s) * c * yz + (-s) * (-s) * yy + c * c * zz;
  o[3] = c * xy + s * xz;
  o[4] = (-s) * xy + c * xz;
  o[5] = (s * (-s) + c * c) * yz + c * (-s) * yy + s * c * zz;
}

// TD coordinates -> global: T' = A^T T A with A rows (vn, vs, vd), the full
// reference formula (src/TDstressFS_C.cpp:28-38).
PQ3D_HD inline void glob_rot(const double* rec, const double T[6], double o[6]) {
  double A[3][3];
  for (int k = 0; k < 3; ++k) { A[0][k] = rec[R_VN + k]; A[1][k] = rec[R_VS + k]; A[2][k] = rec[R_VD + k]; }
  const double Txx = T[0], Tyy = T[1], Tzz = T[2], Txy = T[3], Txz = T[4], Tyz = T[5];
  o[0] = A[0][0] * A[0][0] * Txx + 2 * A[0][0] * A[1][0] * Txy + 2 * A[0][0] * A[2][0] * Txz + 2 * A[1][0] * A[2][0] * Tyz + A[1][0] * A[1][0] * Tyy + A[2][0] * A[2][0] * Tzz;
  o[1] = A[0][1] * A[0][1] * Txx + 2 * A[0][1] * A[1][1] * Txy + 2 * A[0][1] * A[2][1] * Txz + 2 * A[1][1] * A[2][1] * Tyz + A[1][1] * A[1][1] * Tyy + A[2][1] * A[2][1] * Tzz;
  o[2] = A[0][2] * A[0][2] * Txx + 2 * A[0][2] * A[1][2] * Txy + 2 * A[0][2] * A[2][2] * Txz + 2 * A[1][2] * A[2][2] * Tyz + A[1][2] * A[1][2] * Tyy + A[2][2] * A[2][2] * Tzz;
  o[3] = A[0][0] * A[0][1] * Txx + (A[0][0] * A[1][1] + A[0][1] * A[1][0]) * Txy + (A[0][0] * A[2][1] + A[0][1] * A[2][0]) * Txz + (A[2][1] * A[1][0] + A[2][0] * A[1][1]) * Tyz + A[1][1] * A[1][0] * Tyy + A[2][0] * A[2][1] * Tzz;
  o[4] = A[0][0] * A[0][2] * Txx + (A[0][0] * A[1][2] + A[0][2] * A[1][0]) * Txy + (A[0][0] * A[2][2] + A[0][2] * A[2][0]) * Txz + (A[2][2] * A[1][0] + A[2][0] * A[1][2]) * Tyz + A[1][2] * A[1][0] * Tyy + A[2][0] * A[2][2] * Tzz;
  o[5] = A[0][1] * A[0][2] * Txx + (A[0][2] * A[1][1] + A[0][1] * A[1][2]) * Txy + (A[0][2] * A[2][1] + A[0][1] * A[2][2]) * Txz + (A[2][1] * A[1][2] + A[2][2] * A[1][1]) * Tyz + A[1][1] * A[1][2] * Tyy + A[2][1] * A[2][2] * Tzz;
}

PQ3D_HD inline void hooke(const double e[6], double mu, double lam, double s[6]) {
  const double tr = e[0] + e[1] + e[2];
  s[0] = 2 * mu * e[0] + lam * tr;
  s[1] = 2 * mu * e[1] + lam * tr;
  s[2] = 2 * mu * e[2] + lam * tr;
  s[3] = 2 * mu * e[3];
  s[4] = 2 * mu * e[4];
  s[5] = 2 * mu * e[5];
}

// Point in TD coordinates (src/TDstressFS_C.cpp:345-356).
PQ3D_HD inline void to_td(const double* rec, double x, double y, double z, double q[3]) {
  const double d0 = x - rec[R_P2], d1 = y - rec[R_P2 + 1], d2 = z - rec[R_P2 + 2];
  q[0] = rec[R_VN] * d0 + rec[R_VN + 1] * d1 + rec[R_VN + 2] * d2;
  q[1] = rec[R_VS] * d0 + rec[R_VS + 1] * d1 + rec[R_VS + 2] * d2;
  q[2] = rec[R_VD] * d0 + rec[R_VD + 1] * d1 + rec[R_VD + 2] * d2;
}

PQ3D_HD inline void vertex_yz(const double* rec, int k, double& vy, double& vz) {
  if (k == 0) { vy = rec[R_P1]; vz = rec[R_P1 + 1]; }
  else if (k == 1) { vy = 0.0; vz = 0.0; }
  else { vy = rec[R_P3]; vz = rec[R_P3 + 1]; }
}

// Angular-dislocation pair k of the TD seen from q, in its side frame
// (TDSetupS, src/TDstressFS_C.cpp:185-256): side direction (s, c) already
// carries the configuration sign.
PQ3D_HD inline void ad_pair_brackets(const double* rec, const double q[3], int k, double s,
                                     double c, double nu, bool with_bx, AdBrackets& o) {
  double vy, vz;
  vertex_yz(rec, k, vy, vz);
  const double ry = q[1] - vy, rz = q[2] - vz;
  // (y, z) * [[c, -s], [s, c]]^T  ->  y' = c*ry + (-s)*rz, z' = s*ry + c*rz
  const double y1 = c * ry + (-s) * rz;
  const double z1 = s * ry + c * rz;
  ad_brackets(q[0], y1, z1, rec[R_CS + 2 * k], rec[R_CS + 2 * k + 1], nu, with_bx, o);
}

// General Burgers vector (bx=Ts, by=Ss, bz=Ds): global stress and strain at
// one point (src/TDstressFS_C.cpp:274-521 for n=1).
PQ3D_HD inline void tde_point(const double* rec, double x, double y, double z, double Ss,
                              double Ds, double Ts, double mu, double lam, double nu,
                              double stress[6], double strain[6]) {
  double q[3];
  to_td(rec, x, y, z, q);
  const int mode = tde_mode(q[0], q[1], q[2], rec);
  const bool wbx = (Ts != 0.0);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  if (mode != 0) {
    const double sg = (double)mode;  // +1: (-e13, e12, e23); -1: the opposite sides
    double part[3][6];
    for (int k = 0; k < 3; ++k) {
      const double s = sg * rec[R_SV + 2 * k], c = sg * rec[R_SV + 2 * k + 1];
      AdBrackets o;
      ad_pair_brackets(rec, q, k, s, c, nu, wbx, o);
      // slip in the side frame: (by', bz') = A (by, bz), A = [[c, -s], [s, c]]
      const double by1 = c * Ss + (-s) * Ds;
      const double bz1 = s * Ss + c * Ds;
      double e[6];
      ad_combine(o, Ts, by1, bz1, nu, wbx, e);
      side_rot(e, c, s, part[k]);
    }
    for (int j = 0; j < 6; ++j) acc[j] = part[0][j] + part[1][j] + part[2][j];
  }
  glob_rot(rec, acc, strain);
  hooke(strain, mu, lam, stress);
}

// Traction on the receiver plane (frame f = ev1, ev2, ev3):
// (Tt1, Tt2, Tn) = ((sigma.ev3).ev1, (sigma.ev3).ev2, (sigma.ev3).ev3)
// (src/Hmatrix.py:295-313).
PQ3D_HD inline void project(const double s[6], const double* f, double out[3]) {
  const double t1 = s[0] * f[6] + s[3] * f[7] + s[4] * f[8];
  const double t2 = s[3] * f[6] + s[1] * f[7] + s[5] * f[8];
  const double tn = s[4] * f[6] + s[5] * f[7] + s[2] * f[8];
  out[0] = t1 * f[0] + t2 * f[1] + tn * f[2];
  out[1] = t1 * f[3] + t2 * f[4] + tn * f[5];
  out[2] = t1 * f[6] + t2 * f[7] + tn * f[8];
}

// Unit strike (Ss=1) and unit dip (Ds=1) slip of source `rec` at receiver
// point p with frame f: ts = (A1s, A2s, Bs), td = (A1d, A2d, Bd). Bitwise
// equal to two separate tde_point calls followed by project().
PQ3D_HD inline void tde_pair_traction_sd(const double* rec, const double* p, const double* f,
                                         double mu, double lam, double nu, double ts[3],
                                         double td[3]) {
  double q[3];
  to_td(rec, p[0], p[1], p[2], q);
  const int mode = tde_mode(q[0], q[1], q[2], rec);
  double as[6] = {0, 0, 0, 0, 0, 0};
  double ad[6] = {0, 0, 0, 0, 0, 0};
  if (mode != 0) {
    const double sg = (double)mode;
    double ps[3][6], pd[3][6];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double s = sg * rec[R_SV + 2 * k], c = sg * rec[R_SV + 2 * k + 1];
      AdBrackets o;
      ad_pair_brackets(rec, q, k, s, c, nu, false, o);
      double e[6];
      // Ss=1: (by', bz') = (c*1 + (-s)*0, s*1 + c*0) = (c, s)
      ad_combine(o, 0.0, c * 1.0 + (-s) * 0.0, s * 1.0 + c * 0.0, nu, false, e);
      side_rot(e, c, s, ps[k]);
      // Ds=1: (by', bz') = (c*0 + (-s)*1, s*0 + c*1) = (-s, c)
      ad_combine(o, 0.0, c * 0.0 + (-s) * 1.0, s * 0.0 + c * 1.0, nu, false, e);
      side_rot(e, c, s, pd[k]);
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      as[j] = ps[0][j] + ps[1][j] + ps[2][j];
      ad[j] = pd[0][j] + pd[1][j] + pd[2][j];
    }
  }
  double g[6], st[6];
  glob_rot(rec, as, g);
  hooke(g, mu, lam, st);
  project(st, f, ts);
  glob_rot(rec, ad, g);
  hooke(g, mu, lam, st);
  project(st, f, td);
}

}  // namespace pq3d
