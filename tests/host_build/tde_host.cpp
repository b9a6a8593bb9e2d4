// Host (CPU) build of the device TDE math in pyquake3d_amd/csrc/tde_device.h,
// compiled by tests/test_tde_host_build.py only. It lets the CPU test suite
// check, bit for bit, the arithmetic sequence the GPU kernels execute.
#include <cmath>
#include <cstdint>

#include "tde_device.h"

using namespace pq3d;

extern "C" void host_tde_points(const double* pts, int64_t n, const double* tri9, double Ss,
                                double Ds, double Ts, double mu, double lam, double* stress,
                                double* strain) {
  double rec[kSrcRec];
  tde_source_record_host(tri9, tri9 + 3, tri9 + 6, rec);
  const double nu = 1.0 / (1.0 + lam / mu) / 2.0;
  for (int64_t i = 0; i < n; ++i)
    tde_point(rec, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], Ss, Ds, Ts, mu, lam, nu,
              stress + 6 * i, strain + 6 * i);
}

extern "C" void host_tde_block(const double* recv, int64_t m, const double* tri, int64_t n,
                               double mu, double lam, double* out) {
  const double nu = 1.0 / (1.0 + lam / mu) / 2.0;
  const int64_t plane = m * n;
  for (int64_t j = 0; j < n; ++j) {
    double rec[kSrcRec];
    tde_source_record_host(tri + 9 * j, tri + 9 * j + 3, tri + 9 * j + 6, rec);
    for (int64_t i = 0; i < m; ++i) {
      double ts[3], td[3];
      tde_pair_traction_sd(rec, recv + 12 * i, recv + 12 * i + 3, mu, lam, nu, ts, td);
      for (int k = 0; k < 3; ++k) {
        out[k * plane + i * n + j] = ts[k];
        out[(3 + k) * plane + i * n + j] = td[k];
      }
    }
  }
}

#include "tde_hs_device.h"

extern "C" void host_tde_points_hs(const double* pts, int64_t n, const double* tri9, double Ss,
                                   double Ds, double Ts, double mu, double lam, double* stress,
                                   double* strain) {
  double rec[kSrcRec], h[kHsRec];
  tde_source_record_host(tri9, tri9 + 3, tri9 + 6, rec);
  tde_hs_record_host(tri9, tri9 + 3, tri9 + 6, h);
  const double nu = 1.0 / (1.0 + lam / mu) / 2.0;
  for (int64_t i = 0; i < n; ++i)
    tde_point_hs(rec, h, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], Ss, Ds, Ts, mu, lam, nu,
                 stress + 6 * i, strain + 6 * i);
}

extern "C" void host_tde_block_hs(const double* recv, int64_t m, const double* tri, int64_t n,
                                  double mu, double lam, double* out) {
  const double nu = 1.0 / (1.0 + lam / mu) / 2.0;
  const int64_t plane = m * n;
  for (int64_t j = 0; j < n; ++j) {
    double rec[kSrcRec], h[kHsRec];
    tde_source_record_host(tri + 9 * j, tri + 9 * j + 3, tri + 9 * j + 6, rec);
    tde_hs_record_host(tri + 9 * j, tri + 9 * j + 3, tri + 9 * j + 6, h);
    for (int64_t i = 0; i < m; ++i) {
      double ts[3], td[3];
      tde_pair_traction_sd_hs(rec, h, recv + 12 * i, recv + 12 * i + 3, mu, lam, nu, ts, td);
      for (int k = 0; k < 3; ++k) {
        out[k * plane + i * n + j] = ts[k];
        out[(3 + k) * plane + i * n + j] = td[k];
      }
    }
  }
}

#include "tde_disp_device.h"

extern "C" void host_tde_disp_points(const double* pts, int64_t n, const double* tri9, double Ss,
                                     double Ds, double Ts, double nu, int hs, double* disp) {
  double rec[kSrcRec], h[kHsRec];
  tde_source_record_host(tri9, tri9 + 3, tri9 + 6, rec);
  if (hs) tde_hs_record_host(tri9, tri9 + 3, tri9 + 6, h);
  for (int64_t i = 0; i < n; ++i) {
    if (hs)
      tde_disp_point_hs(rec, h, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], Ss, Ds, Ts, nu,
                        disp + 3 * i);
    else
      tde_disp_point(rec, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], Ss, Ds, Ts, nu,
                     disp + 3 * i);
  }
}
