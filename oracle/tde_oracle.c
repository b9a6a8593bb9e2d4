/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, the smoke check and the CPU-baseline leg of the benchmark load
 * this library, and only as the checker. The product path (the HIP library
 * under pyquake3d_amd/) never calls it.
 *
 * A plain-C restatement, point by point, of the reference's full-space
 * triangular-dislocation (TD) stress kernel (Nikkhoo & Walter 2015,
 * "artefact-free" form) and of the traction projection used by the reference's
 * H-matrix evaluators:
 *
 *   ang_dis_strain     follows src/TDstressFS_C.cpp:51-121   (AngDisStrain)
 *   td_mode            follows src/TDstressFS_C.cpp:125-166  (trimodefinder)
 *   td_setup_s         follows src/TDstressFS_C.cpp:185-256  (TDSetupS)
 *   tens_trans         follows src/TDstressFS_C.cpp:21-49    (TensTrans)
 *   oracle_tde_fs_point follows src/TDstressFS_C.cpp:274-521 (TDstressFS, n=1)
 *   oracle_TDstressFS_C follows src/TDstressFS_C.cpp:2048-2095
 *   oracle_TDstressEachSourceAtReceiver_C follows src/TDstressFS_C.cpp:2098-2143
 *   project_traction   follows src/Hmatrix.py:295-313        (GetTtstress)
 *   oracle_traction_block follows src/Hmatrix.py:425-511 (column evaluator,
 *                      unit strike slip then unit dip slip) composed with the
 *                      projection.
 *
 * Pinned by: the golden vectors produced by the reference's own C++ compiled
 * from source (tests/golden/), and by invariants that hold independently of the
 * reference (linearity in the Burgers vector, additivity over a 4-way split of
 * the source triangle, frame covariance, equilibrium div(sigma)=0, far-field
 * point-source limit) -- see tests/test_oracle_tde.py.
 *
 * Mode-0 receivers (on a triangle side, in the triangle plane) give zero
 * strain, as in the reference (src/TDstressFS_C.cpp:446-451). A horizontal
 * source triangle divides by a zero norm exactly as the reference does
 * (src/TDstressFS_C.cpp:317-325) and yields non-finite output.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>
#include <omp.h>

#define TD_PI 3.141592653589793

/* Strain of one angular dislocation at (x,y,z) in its own frame.
 * Output order: xx, yy, zz, xy, xz, yz. */
static void ang_dis_strain(double x, double y, double z, double alpha,
                           double bx, double by, double bz, double nu,
                           double e[6])
{
    const double cA = cos(alpha), sA = sin(alpha);
    const double eta = y * cA - z * sA;
    const double zeta = y * sA + z * cA;
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double r2 = x2 + y2 + z2;
    const double r = sqrt(r2);
    const double r3 = r2 * r;
    const double rmz = r - z;
    const double rz = r * rmz;
    const double r2z2 = r2 * rmz * rmz;
    const double r3z = r3 * rmz;
    const double W = zeta - r;
    const double W2 = W * W;
    const double Wr = W * r;
    const double W2r = W2 * r;
    const double Wr3 = W * r3;
    const double W2r2 = W2 * r2;
    const double C = (r * cA - z) / Wr;
    const double S = (r * sA - y) / Wr;
    const double rmzeta = r - zeta;

    /* gradients of the Burgers function */
    const double fx = (eta / r / rmzeta - y / r / rmz) / 4.0 / TD_PI;
    const double fy = (x / r / rmz - cA * x / r / rmzeta) / 4.0 / TD_PI;
    const double fz = (sA * x / r / rmzeta) / 4.0 / TD_PI;

    const double k = 1.0 / 8.0 / TD_PI / (1.0 - nu);
    const double kx = bx * k;          /* bx/(8 pi (1-nu)) */
    const double kyx = by * x * k;     /* by x/(8 pi (1-nu)) */
    const double kzx = bz * x * sA * k;
    const double ky = by * k;
    const double kz = bz * sA * k;
    const double n21 = 2.0 * nu + 1.0;

    e[0] = bx * fx
         + kx * (eta / Wr + eta * x2 / W2r2 - eta * x2 / Wr3 + y / rz - x2 * y / r2z2 - x2 * y / r3z)
         - kyx * ((n21 / Wr + x2 / W2r2 - x2 / Wr3) * cA + n21 / rz - x2 / r2z2 - x2 / r3z)
         + kzx * (n21 / Wr + x2 / W2r2 - x2 / Wr3);

    e[1] = by * fy
         + kx * ((1.0 / Wr + S * S - y2 / Wr3) * eta + n21 * y / rz - y2 * y / r2z2 - y2 * y / r3z - 2.0 * nu * cA * S)
         - kyx * (1.0 / rz - y2 / r2z2 - y2 / r3z + (1.0 / Wr + S * S - y2 / Wr3) * cA)
         + kzx * (1.0 / Wr + S * S - y2 / Wr3);

    e[2] = bz * fz
         + kx * (eta / W / r + eta * C * C - eta * z2 / Wr3 + y * z / r3 + 2.0 * nu * sA * C)
         - kyx * ((1.0 / Wr + C * C - z2 / Wr3) * cA + z / r3)
         + kzx * (1.0 / Wr + C * C - z2 / Wr3);

    e[3] = bx * fy / 2.0 + by * fx / 2.0
         - kx * (x * y2 / r2z2 - nu * x / rz + x * y2 / r3z - nu * x * cA / Wr + eta * x * S / Wr + eta * x * y / Wr3)
         + ky * (x2 * y / r2z2 - nu * y / rz + x2 * y / r3z + nu * cA * S + x2 * y * cA / Wr3 + x2 * cA * S / Wr)
         - kz * (nu * S + x2 * S / Wr + x2 * y / Wr3);

    e[4] = bx * fz / 2.0 + bz * fx / 2.0
         - kx * (-x * y / r3 + nu * x * sA / Wr + eta * x * C / Wr + eta * x * z / Wr3)
         + ky * (-x2 / r3 + nu / r + nu * cA * C + x2 * z * cA / Wr3 + x2 * cA * C / Wr)
         - kz * (nu * C + x2 * C / Wr + x2 * z / Wr3);

    e[5] = by * fz / 2.0 + bz * fy / 2.0
         + kx * (y2 / r3 - nu / r - nu * cA * C + nu * sA * S + eta * sA * cA / W2
                 - eta * (y * cA + z * sA) / W2r + eta * y * z / W2r2 - eta * y * z / Wr3)
         - kyx * (y / r3 + sA * cA * cA / W2 - cA * (y * cA + z * sA) / W2r + y * z * cA / W2r2 - y * z * cA / Wr3)
         - kzx * (y * z / Wr3 - sA * cA / W2 + (y * cA + z * sA) / W2r - y * z / W2r2);
}

/* Symmetric tensor rotation T' = A^T T A, A row-major 3x3; t in xx,yy,zz,xy,xz,yz. */
static void tens_trans(double t[6], const double A[3][3])
{
    const double T[3][3] = {{t[0], t[3], t[4]}, {t[3], t[1], t[5]}, {t[4], t[5], t[2]}};
    static const int ia[6] = {0, 1, 2, 0, 0, 1};
    static const int ib[6] = {0, 1, 2, 1, 2, 2};
    double out[6];
    for (int c = 0; c < 6; ++c) {
        const int a = ia[c], b = ib[c];
        double s = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                s += A[i][a] * A[j][b] * T[i][j];
        out[c] = s;
    }
    memcpy(t, out, sizeof(out));
}

/* One angular-dislocation pair at TD vertex v with side direction sv
 * (both in TD coordinates), evaluated at point p (TD coordinates). */
static void td_setup_s(const double p[3], double alpha, double bx, double by, double bz,
                       double nu, const double v[3], const double sv[3], double e[6])
{
    const double c = sv[2], s = sv[1];
    const double dy = p[1] - v[1], dz = p[2] - v[2];
    const double y1 = c * dy - s * dz;
    const double z1 = s * dy + c * dz;
    const double by1 = c * by - s * bz;
    const double bz1 = s * by + c * bz;
    ang_dis_strain(p[0], y1, z1, -TD_PI + alpha, bx, by1, bz1, nu, e);
    const double B[3][3] = {{1.0, 0.0, 0.0}, {0.0, c, -s}, {0.0, s, c}};
    tens_trans(e, B);
}

/* Configuration of a point relative to the TD: +1, -1 or 0 (on a side, in plane). */
static int td_mode(const double p[3], const double p1[3], const double p2[3], const double p3[3])
{
    const double x1 = p1[1], x2 = p2[1], x3 = p3[1];
    const double y1 = p1[2], y2 = p2[2], y3 = p3[2];
    const double den = (x1 - x3) * (y2 - y3) - (x2 - x3) * (y1 - y3);
    const double xp = p[1], yp = p[2];
    const double a = ((xp - x3) * (y2 - y3) - (x2 - x3) * (yp - y3)) / den;
    const double b = ((x1 - x3) * (yp - y3) - (xp - x3) * (y1 - y3)) / den;
    const double c = 1.0 - a - b;
    int mode = 1;
    if ((a <= 0 && b > c && c > a) || (b <= 0 && c > a && a > b) || (c <= 0 && a > b && b > c))
        mode = -1;
    else if ((a == 0 && b >= 0 && c >= 0) || (a >= 0 && b == 0 && c >= 0) || (a >= 0 && b >= 0 && c == 0))
        mode = 0;
    if (mode == 0 && p[0] != 0)
        mode = 1;
    return mode;
}

static void unit3(double v[3])
{
    const double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= n; v[1] /= n; v[2] /= n;
}

/* Full-space stress and strain at one point (x,y,z) from one TD. */
void oracle_tde_fs_point(double x, double y, double z,
                         const double *P1, const double *P2, const double *P3,
                         double Ss, double Ds, double Ts, double mu, double lam,
                         double stress[6], double strain[6])
{
    const double bx = Ts, by = Ss, bz = Ds;
    const double nu = 1.0 / (1.0 + lam / mu) / 2.0;

    double vn[3] = {(P2[1] - P1[1]) * (P3[2] - P1[2]) - (P2[2] - P1[2]) * (P3[1] - P1[1]),
                    (P2[2] - P1[2]) * (P3[0] - P1[0]) - (P2[0] - P1[0]) * (P3[2] - P1[2]),
                    (P2[0] - P1[0]) * (P3[1] - P1[1]) - (P2[1] - P1[1]) * (P3[0] - P1[0])};
    unit3(vn);
    /* strike = eZ x normal; a horizontal TD has zero norm and yields non-finite values */
    double vs[3] = {-vn[1], vn[0], 0.0};
    {
        const double n = sqrt(vs[0] * vs[0] + vs[1] * vs[1] + vs[2] * vs[2]);
        if (n == 0) { vs[0] = 0.0; vs[1] = vn[2]; vs[2] = 0.0; }
        vs[0] /= n; vs[1] /= n; vs[2] /= n;
    }
    const double vd[3] = {vn[1] * vs[2] - vn[2] * vs[1],
                          vn[2] * vs[0] - vn[0] * vs[2],
                          vn[0] * vs[1] - vn[1] * vs[0]};

    /* TD coordinates: origin at P2, axes (normal, strike, dip) */
    const double d[3] = {x - P2[0], y - P2[1], z - P2[2]};
    const double q[3] = {vn[0] * d[0] + vn[1] * d[1] + vn[2] * d[2],
                         vs[0] * d[0] + vs[1] * d[1] + vs[2] * d[2],
                         vd[0] * d[0] + vd[1] * d[1] + vd[2] * d[2]};
    const double d1[3] = {P1[0] - P2[0], P1[1] - P2[1], P1[2] - P2[2]};
    const double d3[3] = {P3[0] - P2[0], P3[1] - P2[1], P3[2] - P2[2]};
    const double p1[3] = {vn[0] * d1[0] + vn[1] * d1[1] + vn[2] * d1[2],
                          vs[0] * d1[0] + vs[1] * d1[1] + vs[2] * d1[2],
                          vd[0] * d1[0] + vd[1] * d1[1] + vd[2] * d1[2]};
    const double p2[3] = {0.0, 0.0, 0.0};
    const double p3[3] = {vn[0] * d3[0] + vn[1] * d3[1] + vn[2] * d3[2],
                          vs[0] * d3[0] + vs[1] * d3[1] + vs[2] * d3[2],
                          vd[0] * d3[0] + vd[1] * d3[1] + vd[2] * d3[2]};

    double e12[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    double e13[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
    double e23[3] = {p3[0] - p2[0], p3[1] - p2[1], p3[2] - p2[2]};
    unit3(e12); unit3(e13); unit3(e23);
    const double angA = acos(e12[0] * e13[0] + e12[1] * e13[1] + e12[2] * e13[2]);
    const double angB = acos((-e12[0]) * e23[0] + (-e12[1]) * e23[1] + (-e12[2]) * e23[2]);
    const double angC = acos(e23[0] * e13[0] + e23[1] * e13[1] + e23[2] * e13[2]);

    const int mode = td_mode(q, p1, p2, p3);
    double e[6] = {0, 0, 0, 0, 0, 0};
    if (mode != 0) {
        const double sg = (double)mode; /* +1: (-e13, e12, e23); -1: (e13, -e12, -e23) */
        const double s13[3] = {-sg * e13[0], -sg * e13[1], -sg * e13[2]};
        const double s12[3] = {sg * e12[0], sg * e12[1], sg * e12[2]};
        const double s23[3] = {sg * e23[0], sg * e23[1], sg * e23[2]};
        double ea[6], eb[6], ec[6];
        td_setup_s(q, angA, bx, by, bz, nu, p1, s13, ea);
        td_setup_s(q, angB, bx, by, bz, nu, p2, s12, eb);
        td_setup_s(q, angC, bx, by, bz, nu, p3, s23, ec);
        for (int c = 0; c < 6; ++c) e[c] = ea[c] + eb[c] + ec[c];
    }
    const double A[3][3] = {{vn[0], vn[1], vn[2]}, {vs[0], vs[1], vs[2]}, {vd[0], vd[1], vd[2]}};
    tens_trans(e, A);

    const double tr = e[0] + e[1] + e[2];
    stress[0] = 2.0 * mu * e[0] + lam * tr;
    stress[1] = 2.0 * mu * e[1] + lam * tr;
    stress[2] = 2.0 * mu * e[2] + lam * tr;
    stress[3] = 2.0 * mu * e[3];
    stress[4] = 2.0 * mu * e[4];
    stress[5] = 2.0 * mu * e[5];
    memcpy(strain, e, sizeof(e));
}

/* n receivers x 1 source; outputs n x 6 row-major. Returns 0, or -1 for the
 * half-space flag (not restated in this oracle). */
int oracle_TDstressFS_C(const double *X, const double *Y, const double *Z, size_t n,
                        const double *P1, const double *P2, const double *P3,
                        double Ss, double Ds, double Ts, double mu, double lam,
                        int halfspace, double *stress, double *strain)
{
    if (halfspace) return -1;
    for (size_t i = 0; i < n; ++i)
        oracle_tde_fs_point(X[i], Y[i], Z[i], P1, P2, P3, Ss, Ds, Ts, mu, lam,
                            stress + 6 * i, strain + 6 * i);
    return 0;
}

/* 1 receiver x n sources; P*_flat are n x 3 row-major. */
int oracle_TDstressEachSourceAtReceiver_C(double x, double y, double z,
                                          const double *P1f, const double *P2f, const double *P3f,
                                          size_t ns, double Ss, double Ds, double Ts,
                                          double mu, double lam, int halfspace,
                                          double *stress, double *strain)
{
    if (halfspace) return -1;
    for (size_t i = 0; i < ns; ++i)
        oracle_tde_fs_point(x, y, z, P1f + 3 * i, P2f + 3 * i, P3f + 3 * i, Ss, Ds, Ts, mu, lam,
                            stress + 6 * i, strain + 6 * i);
    return 0;
}

/* Traction on the receiver plane: t = sigma . ev3, projected on (ev1, ev2, ev3).
 * frame = ev1(3), ev2(3), ev3(3). */
static void project_traction(const double s[6], const double *f, double out[3])
{
    const double n0 = f[6], n1 = f[7], n2 = f[8];
    const double t1 = s[0] * n0 + s[3] * n1 + s[4] * n2;
    const double t2 = s[3] * n0 + s[1] * n1 + s[5] * n2;
    const double tn = s[4] * n0 + s[5] * n1 + s[2] * n2;
    out[0] = t1 * f[0] + t2 * f[1] + tn * f[2];
    out[1] = t1 * f[3] + t2 * f[4] + tn * f[5];
    out[2] = t1 * f[6] + t2 * f[7] + tn * f[8];
}

/*
 * Dense traction block: receivers rows[0..m) (centroids xg, frames ev), sources
 * cols[0..n) (triangles tri_p123, 9 doubles each: P1,P2,P3).
 * out is 6 x m x n row-major, operators in the order A1s, A2s, Bs, A1d, A2d, Bd
 * (strike slip Ss=1 gives A1s,A2s,Bs; dip slip Ds=1 gives A1d,A2d,Bd).
 */
void oracle_traction_block(const double *xg, const double *ev, const double *tri_p123,
                           const int64_t *rows, int64_t m, const int64_t *cols, int64_t n,
                           double mu, double lam, double *out)
{
    const size_t mn = (size_t)m * (size_t)n;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < m; ++i) {
        const int64_t r = rows[i];
        const double *p = xg + 3 * r;
        const double *f = ev + 9 * r;
        for (int64_t j = 0; j < n; ++j) {
            const double *t = tri_p123 + 9 * cols[j];
            double s[6], e[6], tr[3];
            oracle_tde_fs_point(p[0], p[1], p[2], t, t + 3, t + 6, 1.0, 0.0, 0.0, mu, lam, s, e);
            project_traction(s, f, tr);
            out[0 * mn + (size_t)i * n + j] = tr[0];
            out[1 * mn + (size_t)i * n + j] = tr[1];
            out[2 * mn + (size_t)i * n + j] = tr[2];
            oracle_tde_fs_point(p[0], p[1], p[2], t, t + 3, t + 6, 0.0, 1.0, 0.0, mu, lam, s, e);
            project_traction(s, f, tr);
            out[3 * mn + (size_t)i * n + j] = tr[0];
            out[4 * mn + (size_t)i * n + j] = tr[1];
            out[5 * mn + (size_t)i * n + j] = tr[2];
        }
    }
}

/* Threads the OpenMP runtime will use (for reporting the CPU baseline). */
int oracle_num_threads(void)
{
    int n = 1;
#pragma omp parallel
    {
#pragma omp single
        n = omp_get_num_threads();
    }
    return n;
}
