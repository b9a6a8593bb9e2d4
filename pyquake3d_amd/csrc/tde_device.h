// Device-side full-space triangular-dislocation (TD) stress, FP64.
//
// Math: Nikkhoo & Walter (2015), artefact-free full-space solution, in the
// form the reference evaluates it (src/TDstressFS_C.cpp:51-521). The work is
// split MI355X-first:
//   * everything that depends only on the source triangle (frame, vertices in
//     TD coordinates, side directions, the three angular-dislocation angles) is
//     computed once per triangle into a 32-double "source record"
//     (tde_source_record) instead of once per (receiver, source) pair as the
//     reference does (src/TDstressFS_C.cpp:292-395);
//   * the per-pair work (ang_dis_*) uses reciprocals of the shared
//     denominators instead of ~60 FP64 divisions;
//   * the traction assembly evaluates unit strike and unit dip slip together:
//     the angular-dislocation strain is linear in the Burgers vector, so both
//     come out of one pass over the geometry terms (tde_pair_traction_sd).
#pragma once
#include <hip/hip_runtime.h>

namespace pq3d {

constexpr double kPi = 3.141592653589793;
constexpr int kSrcRec = 32;  // doubles per source record (256 B)

// Source record layout (doubles):
//  0..2  normal (vn)      3..5 strike (vs)     6..8 dip (vd)
//  9..11 P2 (origin of TD coordinates)
// 12,13 p1 (y,z) in TD coordinates        14,15 p3 (y,z)
// 16    barycentric denominator
// 17,18 side 0 direction (y,z) = -e13     19,20 side 1 = e12    21,22 side 2 = e23
// 23,24 cos, sin of (-pi + angle A)       25,26 same for B      27,28 same for C
// 29..31 unused
enum : int {
  R_VN = 0, R_VS = 3, R_VD = 6, R_P2 = 9, R_P1 = 12, R_P3 = 14, R_DEN = 16,
  R_SV = 17, R_CS = 23
};

__host__ __device__ inline void unit3(double v[3]) {
  const double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] /= n; v[1] /= n; v[2] /= n;
}

// Per-triangle setup (src/TDstressFS_C.cpp:292-395). A horizontal triangle
// divides by the zero strike norm exactly as the reference does and produces a
// non-finite record, hence non-finite stresses.
__device__ inline void tde_source_record(const double* P1, const double* P2, const double* P3,
                                         double* rec) {
  double vn[3] = {(P2[1] - P1[1]) * (P3[2] - P1[2]) - (P2[2] - P1[2]) * (P3[1] - P1[1]),
                  (P2[2] - P1[2]) * (P3[0] - P1[0]) - (P2[0] - P1[0]) * (P3[2] - P1[2]),
                  (P2[0] - P1[0]) * (P3[1] - P1[1]) - (P2[1] - P1[1]) * (P3[0] - P1[0])};
  unit3(vn);
  double vs[3] = {-vn[1], vn[0], 0.0};
  {
    const double n = sqrt(vs[0] * vs[0] + vs[1] * vs[1]);
    if (n == 0.0) { vs[0] = 0.0; vs[1] = vn[2]; vs[2] = 0.0; }
    vs[0] /= n; vs[1] /= n; vs[2] /= n;
  }
  const double vd[3] = {vn[1] * vs[2] - vn[2] * vs[1], vn[2] * vs[0] - vn[0] * vs[2],
                        vn[0] * vs[1] - vn[1] * vs[0]};
  const double d1[3] = {P1[0] - P2[0], P1[1] - P2[1], P1[2] - P2[2]};
  const double d3[3] = {P3[0] - P2[0], P3[1] - P2[1], P3[2] - P2[2]};
  const double p1[3] = {vn[0] * d1[0] + vn[1] * d1[1] + vn[2] * d1[2],
                        vs[0] * d1[0] + vs[1] * d1[1] + vs[2] * d1[2],
                        vd[0] * d1[0] + vd[1] * d1[1] + vd[2] * d1[2]};
  const double p3[3] = {vn[0] * d3[0] + vn[1] * d3[1] + vn[2] * d3[2],
                        vs[0] * d3[0] + vs[1] * d3[1] + vs[2] * d3[2],
                        vd[0] * d3[0] + vd[1] * d3[1] + vd[2] * d3[2]};
  double e12[3] = {-p1[0], -p1[1], -p1[2]};
  double e13[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
  double e23[3] = {p3[0], p3[1], p3[2]};
  unit3(e12); unit3(e13); unit3(e23);
  const double aA = acos(e12[0] * e13[0] + e12[1] * e13[1] + e12[2] * e13[2]);
  const double aB = acos((-e12[0]) * e23[0] + (-e12[1]) * e23[1] + (-e12[2]) * e23[2]);
  const double aC = acos(e23[0] * e13[0] + e23[1] * e13[1] + e23[2] * e13[2]);

  for (int k = 0; k < 3; ++k) { rec[R_VN + k] = vn[k]; rec[R_VS + k] = vs[k]; rec[R_VD + k] = vd[k]; rec[R_P2 + k] = P2[k]; }
  rec[R_P1] = p1[1]; rec[R_P1 + 1] = p1[2];
  rec[R_P3] = p3[1]; rec[R_P3 + 1] = p3[2];
  // p2 = (0,0,0): (x1-x3)(y2-y3) - (x2-x3)(y1-y3) with x2=y2=0
  rec[R_DEN] = (p1[1] - p3[1]) * (0.0 - p3[2]) - (0.0 - p3[1]) * (p1[2] - p3[2]);
  rec[R_SV + 0] = -e13[1]; rec[R_SV + 1] = -e13[2];
  rec[R_SV + 2] = e12[1];  rec[R_SV + 3] = e12[2];
  rec[R_SV + 4] = e23[1];  rec[R_SV + 5] = e23[2];
  rec[R_CS + 0] = cos(-kPi + aA); rec[R_CS + 1] = sin(-kPi + aA);
  rec[R_CS + 2] = cos(-kPi + aB); rec[R_CS + 3] = sin(-kPi + aB);
  rec[R_CS + 4] = cos(-kPi + aC); rec[R_CS + 5] = sin(-kPi + aC);
  rec[29] = 0.0; rec[30] = 0.0; rec[31] = 0.0;
}

// Configuration of a TD-coordinate point: +1, -1, or 0 (on a side and in the
// TD plane). Same decisions as src/TDstressFS_C.cpp:125-166, including the
// exact comparisons with zero.
__device__ inline int tde_mode(double q0, double q1, double q2, const double* rec) {
#pragma clang fp contract(off)
  const double x1 = rec[R_P1], y1 = rec[R_P1 + 1];
  const double x3 = rec[R_P3], y3 = rec[R_P3 + 1];
  const double x2 = 0.0, y2 = 0.0;
  const double den = rec[R_DEN];
  const double a = ((q1 - x3) * (y2 - y3) - (x2 - x3) * (q2 - y3)) / den;
  const double b = ((x1 - x3) * (q2 - y3) - (q1 - x3) * (y1 - y3)) / den;
  const double c = 1.0 - a - b;
  int mode = 1;
  if ((a <= 0 && b > c && c > a) || (b <= 0 && c > a && a > b) || (c <= 0 && a > b && b > c))
    mode = -1;
  else if ((a == 0 && b >= 0 && c >= 0) || (a >= 0 && b == 0 && c >= 0) ||
           (a >= 0 && b >= 0 && c == 0))
    mode = 0;
  if (mode == 0 && q0 != 0) mode = 1;
  return mode;
}

// Geometry terms of one angular dislocation shared by all Burgers components.
struct AdTerms {
  double x, y, z, cA, sA, eta, x2, y2, z2;
  double ir, ir3, irz, ir2z2, ir3z, iWr, iW2r2, iWr3, iW2, iW2r, C, S;
  double fx, fy, fz;  // Burgers-function gradients
};

__device__ inline void ad_terms(double x, double y, double z, double cA, double sA, AdTerms& t) {
  t.x = x; t.y = y; t.z = z; t.cA = cA; t.sA = sA;
  t.eta = y * cA - z * sA;
  const double zeta = y * sA + z * cA;
  t.x2 = x * x; t.y2 = y * y; t.z2 = z * z;
  const double r2 = t.x2 + t.y2 + t.z2;
  const double r = sqrt(r2);
  const double r3 = r2 * r;
  const double rmz = r - z;
  const double W = zeta - r;
  const double rmzeta = r - zeta;
  t.ir = 1.0 / r;
  t.ir3 = 1.0 / r3;
  t.irz = 1.0 / (r * rmz);
  t.ir2z2 = 1.0 / (r2 * rmz * rmz);
  t.ir3z = 1.0 / (r3 * rmz);
  const double Wr = W * r;
  t.iWr = 1.0 / Wr;
  t.iW2r2 = 1.0 / (W * W * r2);
  t.iWr3 = 1.0 / (W * r3);
  t.iW2 = 1.0 / (W * W);
  t.iW2r = 1.0 / (W * W * r);
  t.C = (r * cA - z) * t.iWr;
  t.S = (r * sA - y) * t.iWr;
  const double irmzeta = 1.0 / rmzeta;
  const double irmz = 1.0 / rmz;
  const double q = 1.0 / (4.0 * kPi);
  t.fx = (t.eta * t.ir * irmzeta - y * t.ir * irmz) * q;
  t.fy = (x * t.ir * irmz - cA * x * t.ir * irmzeta) * q;
  t.fz = (sA * x * t.ir * irmzeta) * q;
}

// Strain of the angular dislocation for Burgers vector (bx, by, bz),
// xx,yy,zz,xy,xz,yz (src/TDstressFS_C.cpp:90-119).
__device__ inline void ad_strain(const AdTerms& t, double bx, double by, double bz, double nu,
                                 double e[6]) {
  const double k = 1.0 / (8.0 * kPi * (1.0 - nu));
  const double x = t.x, y = t.y, z = t.z, cA = t.cA, sA = t.sA, eta = t.eta;
  const double x2 = t.x2, y2 = t.y2, z2 = t.z2, C = t.C, S = t.S;
  const double kx = bx * k, kyx = by * x * k, kzx = bz * x * sA * k, ky = by * k, kz = bz * sA * k;
  const double n21 = 2.0 * nu + 1.0;
  const double P = 1.0 * t.iWr + S * S - y2 * t.iWr3;  // shared yy bracket
  const double Q = 1.0 * t.iWr + C * C - z2 * t.iWr3;  // shared zz bracket
  const double R = n21 * t.iWr + x2 * t.iW2r2 - x2 * t.iWr3;
  const double ycz = y * cA + z * sA;
  e[0] = bx * t.fx
       + kx * (eta * t.iWr + eta * x2 * t.iW2r2 - eta * x2 * t.iWr3 + y * t.irz - x2 * y * t.ir2z2 - x2 * y * t.ir3z)
       - kyx * (R * cA + n21 * t.irz - x2 * t.ir2z2 - x2 * t.ir3z)
       + kzx * R;
  e[1] = by * t.fy
       + kx * (P * eta + n21 * y * t.irz - y2 * y * t.ir2z2 - y2 * y * t.ir3z - 2.0 * nu * cA * S)
       - kyx * (t.irz - y2 * t.ir2z2 - y2 * t.ir3z + P * cA)
       + kzx * P;
  e[2] = bz * t.fz
       + kx * (eta * t.iWr + eta * C * C - eta * z2 * t.iWr3 + y * z * t.ir3 + 2.0 * nu * sA * C)
       - kyx * (Q * cA + z * t.ir3)
       + kzx * Q;
  e[3] = 0.5 * bx * t.fy + 0.5 * by * t.fx
       - kx * (x * y2 * t.ir2z2 - nu * x * t.irz + x * y2 * t.ir3z - nu * x * cA * t.iWr + eta * x * S * t.iWr + eta * x * y * t.iWr3)
       + ky * (x2 * y * t.ir2z2 - nu * y * t.irz + x2 * y * t.ir3z + nu * cA * S + x2 * y * cA * t.iWr3 + x2 * cA * S * t.iWr)
       - kz * (nu * S + x2 * S * t.iWr + x2 * y * t.iWr3);
  e[4] = 0.5 * bx * t.fz + 0.5 * bz * t.fx
       - kx * (-x * y * t.ir3 + nu * x * sA * t.iWr + eta * x * C * t.iWr + eta * x * z * t.iWr3)
       + ky * (-x2 * t.ir3 + nu * t.ir + nu * cA * C + x2 * z * cA * t.iWr3 + x2 * cA * C * t.iWr)
       - kz * (nu * C + x2 * C * t.iWr + x2 * z * t.iWr3);
  e[5] = 0.5 * by * t.fz + 0.5 * bz * t.fy
       + kx * (y2 * t.ir3 - nu * t.ir - nu * cA * C + nu * sA * S + eta * sA * cA * t.iW2
               - eta * ycz * t.iW2r + eta * y * z * t.iW2r2 - eta * y * z * t.iWr3)
       - kyx * (y * t.ir3 + sA * cA * cA * t.iW2 - cA * ycz * t.iW2r + y * z * cA * t.iW2r2 - y * z * cA * t.iWr3)
       - kzx * (y * z * t.iWr3 - sA * cA * t.iW2 + ycz * t.iW2r - y * z * t.iW2r2);
}

// Unit-Burgers strains with bx = 0: ey for (by,bz)=(1,0), ez for (0,1).
__device__ inline void ad_strain_yz(const AdTerms& t, double nu, double ey[6], double ez[6]) {
  const double k = 1.0 / (8.0 * kPi * (1.0 - nu));
  const double x = t.x, y = t.y, z = t.z, cA = t.cA, sA = t.sA;
  const double x2 = t.x2, y2 = t.y2, z2 = t.z2, C = t.C, S = t.S;
  const double kyx = x * k, kzx = x * sA * k, kz = sA * k;
  const double n21 = 2.0 * nu + 1.0;
  const double P = t.iWr + S * S - y2 * t.iWr3;
  const double Q = t.iWr + C * C - z2 * t.iWr3;
  const double R = n21 * t.iWr + x2 * t.iW2r2 - x2 * t.iWr3;
  const double ycz = y * cA + z * sA;
  // by = 1
  ey[0] = -kyx * (R * cA + n21 * t.irz - x2 * t.ir2z2 - x2 * t.ir3z);
  ey[1] = t.fy - kyx * (t.irz - y2 * t.ir2z2 - y2 * t.ir3z + P * cA);
  ey[2] = -kyx * (Q * cA + z * t.ir3);
  ey[3] = 0.5 * t.fx
        + k * (x2 * y * t.ir2z2 - nu * y * t.irz + x2 * y * t.ir3z + nu * cA * S + x2 * y * cA * t.iWr3 + x2 * cA * S * t.iWr);
  ey[4] = k * (-x2 * t.ir3 + nu * t.ir + nu * cA * C + x2 * z * cA * t.iWr3 + x2 * cA * C * t.iWr);
  ey[5] = 0.5 * t.fz
        - kyx * (y * t.ir3 + sA * cA * cA * t.iW2 - cA * ycz * t.iW2r + y * z * cA * t.iW2r2 - y * z * cA * t.iWr3);
  // bz = 1
  ez[0] = kzx * R;
  ez[1] = kzx * P;
  ez[2] = t.fz + kzx * Q;
  ez[3] = -kz * (nu * S + x2 * S * t.iWr + x2 * y * t.iWr3);
  ez[4] = 0.5 * t.fx - kz * (nu * C + x2 * C * t.iWr + x2 * z * t.iWr3);
  ez[5] = 0.5 * t.fy
        - kzx * (y * z * t.iWr3 - sA * cA * t.iW2 + ycz * t.iW2r - y * z * t.iW2r2);
}

// Rotate a strain given in the (x, y', z') frame of a side (y' = c y - s z,
// z' = s y + c z) back to TD coordinates: T' = B^T T B with
// B = [[1,0,0],[0,c,-s],[0,s,c]] (src/TDstressFS_C.cpp:238-252).
__device__ inline void rot_side_add(const double e[6], double c, double s, double acc[6]) {
  const double xx = e[0], yy = e[1], zz = e[2], xy = e[3], xz = e[4], yz = e[5];
  acc[0] += xx;
  acc[1] += c * c * yy + 2.0 * c * s * yz + s * s * zz;
  acc[2] += s * s * yy - 2.0 * s * c * yz + c * c * zz;
  acc[3] += c * xy + s * xz;
  acc[4] += -s * xy + c * xz;
  acc[5] += -c * s * yy + (c * c - s * s) * yz + s * c * zz;
}

// TD coordinates -> global: T' = A^T T A with A rows (vn, vs, vd).
__device__ inline void rot_global(const double* rec, const double t[6], double o[6]) {
  const double* a0 = rec + R_VN;
  const double* a1 = rec + R_VS;
  const double* a2 = rec + R_VD;
  const double T00 = t[0], T11 = t[1], T22 = t[2], T01 = t[3], T02 = t[4], T12 = t[5];
  // o_ab = sum_ij A_ia A_jb T_ij
#define PQ3D_ROT(a, b)                                                                  \
  (a0[a] * (a0[b] * T00 + a1[b] * T01 + a2[b] * T02) +                                   \
   a1[a] * (a0[b] * T01 + a1[b] * T11 + a2[b] * T12) +                                   \
   a2[a] * (a0[b] * T02 + a1[b] * T12 + a2[b] * T22))
  o[0] = PQ3D_ROT(0, 0);
  o[1] = PQ3D_ROT(1, 1);
  o[2] = PQ3D_ROT(2, 2);
  o[3] = PQ3D_ROT(0, 1);
  o[4] = PQ3D_ROT(0, 2);
  o[5] = PQ3D_ROT(1, 2);
#undef PQ3D_ROT
}

__device__ inline void hooke(const double e[6], double mu, double lam, double s[6]) {
  const double tr = e[0] + e[1] + e[2];
  s[0] = 2.0 * mu * e[0] + lam * tr;
  s[1] = 2.0 * mu * e[1] + lam * tr;
  s[2] = 2.0 * mu * e[2] + lam * tr;
  s[3] = 2.0 * mu * e[3];
  s[4] = 2.0 * mu * e[4];
  s[5] = 2.0 * mu * e[5];
}

// Point (x,y,z) in TD coordinates relative to the record.
__device__ inline void to_td(const double* rec, double x, double y, double z, double q[3]) {
  const double d0 = x - rec[R_P2], d1 = y - rec[R_P2 + 1], d2 = z - rec[R_P2 + 2];
  q[0] = rec[R_VN] * d0 + rec[R_VN + 1] * d1 + rec[R_VN + 2] * d2;
  q[1] = rec[R_VS] * d0 + rec[R_VS + 1] * d1 + rec[R_VS + 2] * d2;
  q[2] = rec[R_VD] * d0 + rec[R_VD + 1] * d1 + rec[R_VD + 2] * d2;
}

__device__ inline void vertex_yz(const double* rec, int k, double& vy, double& vz) {
  if (k == 0) { vy = rec[R_P1]; vz = rec[R_P1 + 1]; }
  else if (k == 1) { vy = 0.0; vz = 0.0; }
  else { vy = rec[R_P3]; vz = rec[R_P3 + 1]; }
}

// General Burgers vector (bx=Ts, by=Ss, bz=Ds): global strain and stress at one
// point (src/TDstressFS_C.cpp:274-521 for n=1).
__device__ inline void tde_point(const double* rec, double x, double y, double z, double Ss,
                                 double Ds, double Ts, double mu, double lam, double nu,
                                 double stress[6], double strain[6]) {
  double q[3];
  to_td(rec, x, y, z, q);
  const int mode = tde_mode(q[0], q[1], q[2], rec);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  if (mode != 0) {
    const double sg = (double)mode;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double s = sg * rec[R_SV + 2 * k], c = sg * rec[R_SV + 2 * k + 1];
      double vy, vz;
      vertex_yz(rec, k, vy, vz);
      const double dy = q[1] - vy, dz = q[2] - vz;
      AdTerms t;
      ad_terms(q[0], c * dy - s * dz, s * dy + c * dz, rec[R_CS + 2 * k], rec[R_CS + 2 * k + 1], t);
      double e[6];
      ad_strain(t, Ts, c * Ss - s * Ds, s * Ss + c * Ds, nu, e);
      rot_side_add(e, c, s, acc);
    }
  }
  rot_global(rec, acc, strain);
  hooke(strain, mu, lam, stress);
}

// Traction on a receiver plane (frame f = ev1, ev2, ev3) from stress s:
// (Tt1, Tt2, Tn) (src/Hmatrix.py:295-313).
__device__ inline void project(const double s[6], const double* f, double out[3]) {
  const double t1 = s[0] * f[6] + s[3] * f[7] + s[4] * f[8];
  const double t2 = s[3] * f[6] + s[1] * f[7] + s[5] * f[8];
  const double tn = s[4] * f[6] + s[5] * f[7] + s[2] * f[8];
  out[0] = t1 * f[0] + t2 * f[1] + tn * f[2];
  out[1] = t1 * f[3] + t2 * f[4] + tn * f[5];
  out[2] = t1 * f[6] + t2 * f[7] + tn * f[8];
}

// Unit strike (Ss=1) and unit dip (Ds=1) slip on source `rec` seen at receiver
// point p with frame f: ts = (A1s, A2s, Bs), td = (A1d, A2d, Bd).
__device__ inline void tde_pair_traction_sd(const double* rec, const double* p, const double* f,
                                            double mu, double lam, double nu, double ts[3],
                                            double td[3]) {
  double q[3];
  to_td(rec, p[0], p[1], p[2], q);
  const int mode = tde_mode(q[0], q[1], q[2], rec);
  double as[6] = {0, 0, 0, 0, 0, 0};
  double ad[6] = {0, 0, 0, 0, 0, 0};
  if (mode != 0) {
    const double sg = (double)mode;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double s = sg * rec[R_SV + 2 * k], c = sg * rec[R_SV + 2 * k + 1];
      double vy, vz;
      vertex_yz(rec, k, vy, vz);
      const double dy = q[1] - vy, dz = q[2] - vz;
      AdTerms t;
      ad_terms(q[0], c * dy - s * dz, s * dy + c * dz, rec[R_CS + 2 * k], rec[R_CS + 2 * k + 1], t);
      double ey[6], ez[6], es[6], ed[6];
      ad_strain_yz(t, nu, ey, ez);
      // Ss=1 -> (by', bz') = (c, s); Ds=1 -> (-s, c)
#pragma unroll
      for (int j = 0; j < 6; ++j) { es[j] = c * ey[j] + s * ez[j]; ed[j] = c * ez[j] - s * ey[j]; }
      rot_side_add(es, c, s, as);
      rot_side_add(ed, c, s, ad);
    }
  }
  double g[6], st[6];
  rot_global(rec, as, g);
  hooke(g, mu, lam, st);
  project(st, f, ts);
  rot_global(rec, ad, g);
  hooke(g, mu, lam, st);
  project(st, f, td);
}

}  // namespace pq3d
