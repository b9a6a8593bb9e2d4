// Driver of the sanitizer build (tests/test_sanitized_host_build.py): runs
// the host build of the device TDE arithmetic (tde_host.cpp) and the plain-C
// oracle (oracle/tde_oracle.c) over the records of an input file and writes
// their outputs. Input: int64 n, int64 halfspace flag, then n records of 17
// doubles (point 3, triangle 9, Burgers vector Ss Ds Ts, mu, lambda).
// Output: n rows of 15 doubles (device-code stress 6, oracle stress 6 --
// NaN in the half space, which the C oracle does not restate -- and the
// device-code displacement 3).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

extern "C" {
void host_tde_points(const double* pts, int64_t n, const double* tri9, double Ss, double Ds,
                     double Ts, double mu, double lam, double* stress, double* strain);
void host_tde_points_hs(const double* pts, int64_t n, const double* tri9, double Ss, double Ds,
                        double Ts, double mu, double lam, double* stress, double* strain);
void host_tde_disp_points(const double* pts, int64_t n, const double* tri9, double Ss,
                          double Ds, double Ts, double nu, int hs, double* disp);
int oracle_TDstressFS_C(const double* X, const double* Y, const double* Z, size_t n,
                        const double* P1, const double* P2, const double* P3, double Ss,
                        double Ds, double Ts, double mu, double lam, int halfspace,
                        double* stress, double* strain);
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s input output\n", argv[0]);
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 3;
  int64_t hdr[2];
  if (std::fread(hdr, sizeof(int64_t), 2, f) != 2) return 4;
  const int64_t n = hdr[0];
  const int hs = (int)hdr[1];
  std::vector<double> rec((size_t)n * 17);
  if (std::fread(rec.data(), sizeof(double), rec.size(), f) != rec.size()) return 5;
  std::fclose(f);
  std::vector<double> out((size_t)n * 15);
  for (int64_t i = 0; i < n; ++i) {
    const double* r = rec.data() + 17 * i;
    const double *pt = r, *tri = r + 3, *bv = r + 12, mu = r[15], lam = r[16];
    double* o = out.data() + 15 * i;
    double strain[6];
    if (hs)
      host_tde_points_hs(pt, 1, tri, bv[0], bv[1], bv[2], mu, lam, o, strain);
    else
      host_tde_points(pt, 1, tri, bv[0], bv[1], bv[2], mu, lam, o, strain);
    if (hs) {
      for (int k = 0; k < 6; ++k) o[6 + k] = NAN;
    } else {
      oracle_TDstressFS_C(pt, pt + 1, pt + 2, 1, tri, tri + 3, tri + 6, bv[0], bv[1], bv[2], mu,
                          lam, 0, o + 6, strain);
    }
    const double nu = 1.0 / (1.0 + lam / mu) / 2.0;
    host_tde_disp_points(pt, 1, tri, bv[0], bv[1], bv[2], nu, hs, o + 12);
  }
  f = std::fopen(argv[2], "wb");
  if (!f) return 6;
  std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
  return 0;
}
