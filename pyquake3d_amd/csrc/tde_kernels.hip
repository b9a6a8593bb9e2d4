// FP64 triangular-dislocation kernels and their C-ABI.
//
//  * pq3d_tde_source_records   per-triangle setup on the host (C library trig)
//  * pq3d_tde_stress_points    n receivers x 1 source (device pointers)
//  * pq3d_tde_stress_sources   1 receiver x n sources (device pointers)
//  * pq3d_tde_traction_block   m receivers x n sources, unit strike and unit dip
//                              slip, projected on the receiver frames -> the six
//                              traction operators A1s, A2s, Bs, A1d, A2d, Bd
//  * TDstressFS_C / TDstressEachSourceAtReceiver_C
//                              drop-in host-pointer entry points with the
//                              reference's signatures (src/TDstressFS_C.cpp:2048,
//                              :2098) built on the kernels above.
//  * *_hs variants             the same for the half space (free surface z = 0,
//                              src/TDstressFS_C.cpp:1911-1999), with a second
//                              per-triangle record (pq3d_tde_hs_records).
//  * pq3d_tde_disp_points / pq3d_tde_disp_sum
//                              displacement (src/DH_greenfunction.py TDdispFS,
//                              TDdispHS): n points x 1 triangle, and points x
//                              triangles summed with per-triangle slip.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "pq3d_common.h"
#include "tde_device.h"
#include "tde_hs_device.h"
#include "tde_disp_device.h"

namespace pq3d {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
void clear_error() { g_err[0] = 0; }

// ---------------------------------------------------------------- kernels

__device__ inline void load_rec(const double* __restrict__ src, double* r) {
#pragma unroll
  for (int k = 0; k < kSrcRec; ++k) r[k] = src[k];
}

// Reference symbol 1: n receivers (X,Y,Z) x one source record; with HS, the
// half-space stress (hrec: the triangle's half-space record).
template <bool HS>
__global__ void __launch_bounds__(256) k_stress_points(const double* __restrict__ X,
                                                       const double* __restrict__ Y,
                                                       const double* __restrict__ Z, int64_t n,
                                                       const double* __restrict__ rec1,
                                                       const double* __restrict__ hrec1,
                                                       double Ss, double Ds, double Ts, double mu,
                                                       double lam, double nu,
                                                       double* __restrict__ stress,
                                                       double* __restrict__ strain) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r[kSrcRec];
  load_rec(rec1, r);
  double s[6], e[6];
  if constexpr (HS)
    tde_point_hs(r, hrec1, X[i], Y[i], Z[i], Ss, Ds, Ts, mu, lam, nu, s, e);
  else
    tde_point(r, X[i], Y[i], Z[i], Ss, Ds, Ts, mu, lam, nu, s, e);
#pragma unroll
  for (int k = 0; k < 6; ++k) { stress[6 * i + k] = s[k]; strain[6 * i + k] = e[k]; }
}

// Reference symbol 2: one receiver x n source records (HS: and their
// half-space records).
template <bool HS>
__global__ void __launch_bounds__(256) k_stress_sources(double x, double y, double z,
                                                        const double* __restrict__ rec,
                                                        const double* __restrict__ hrec,
                                                        int64_t ns, double Ss, double Ds,
                                                        double Ts, double mu, double lam,
                                                        double nu, double* __restrict__ stress,
                                                        double* __restrict__ strain) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  double r[kSrcRec];
  load_rec(rec + (int64_t)kSrcRec * i, r);
  double s[6], e[6];
  if constexpr (HS)
    tde_point_hs(r, hrec + (int64_t)kHsRec * i, x, y, z, Ss, Ds, Ts, mu, lam, nu, s, e);
  else
    tde_point(r, x, y, z, Ss, Ds, Ts, mu, lam, nu, s, e);
#pragma unroll
  for (int k = 0; k < 6; ++k) { stress[6 * i + k] = s[k]; strain[6 * i + k] = e[k]; }
}

constexpr int kTbCols = 256;  // sources per workgroup (one per lane)
constexpr int kTbRows = 8;    // receivers per thread

// Traction block: lane <-> source column (coalesced stores along j), each
// thread walks kTbRows receivers whose 12 doubles (point + frame) are uniform
// across the workgroup. Output: out[op * plane + i * si + j * sj], op order
// A1s, A2s, Bs, A1d, A2d, Bd. rows/cols may be null (identity + offset).
// Each lane's 32-double source record sits in LDS (row stride kSrcRec + 1
// doubles, so the 64 lanes of a read hit distinct banks) instead of 64
// VGPRs: the full-space kernel then fits 256 registers and runs two waves
// per SIMD (52.3 -> 44.9 ms for N = 20,000). The half-space kernel stays at
// one wave per SIMD: capped at 256 registers it spills ~300 VGPRs.
template <typename OutT, bool HS>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, kTbCols),
                               amdgpu_waves_per_eu(HS ? 1 : 2))) k_traction_block(
    const double* __restrict__ recv, const double* __restrict__ rec,
    const double* __restrict__ hrec,
    const int64_t* __restrict__ rows, int64_t row0, int64_t m, const int64_t* __restrict__ cols,
    int64_t col0, int64_t n, double mu, double lam, double nu, OutT* __restrict__ out, int64_t si,
    int64_t sj, int64_t plane) {
  __shared__ double recs[kTbCols * (kSrcRec + 1)];
  const int64_t j = (int64_t)blockIdx.x * kTbCols + threadIdx.x;
  const int64_t ib = (int64_t)blockIdx.y * kTbRows;
  if (j >= n) return;
  const int64_t src = cols ? cols[j] : col0 + j;
  // only this lane reads its slot: no barrier
  double* r = recs + threadIdx.x * (kSrcRec + 1);
  load_rec(rec + (int64_t)kSrcRec * src, r);
  const int64_t iend = (ib + kTbRows < m) ? ib + kTbRows : m;
  for (int64_t i = ib; i < iend; ++i) {
    const int64_t rcv = rows ? rows[i] : row0 + i;
    const double* p = recv + 12 * rcv;
    double ts[3], td[3];
    if constexpr (HS)
      tde_pair_traction_sd_hs(r, hrec + (int64_t)kHsRec * src, p, p + 3, mu, lam, nu, ts, td);
    else
      tde_pair_traction_sd(r, p, p + 3, mu, lam, nu, ts, td);
    OutT* o = out + i * si + j * sj;
    o[0 * plane] = (OutT)ts[0];
    o[1 * plane] = (OutT)ts[1];
    o[2 * plane] = (OutT)ts[2];
    o[3 * plane] = (OutT)td[0];
    o[4 * plane] = (OutT)td[1];
    o[5 * plane] = (OutT)td[2];
  }
}

// Displacement of one triangle at n points (TDdispFS / TDdispHS of
// src/DH_greenfunction.py:99, :402); out n x 3 row-major (east, north, up).
template <bool HS>
__global__ void __launch_bounds__(256) k_disp_points(const double* __restrict__ X,
                                                     const double* __restrict__ Y,
                                                     const double* __restrict__ Z, int64_t n,
                                                     const double* __restrict__ rec1,
                                                     const double* __restrict__ hrec1, double Ss,
                                                     double Ds, double Ts, double nu,
                                                     double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r[kSrcRec];
  load_rec(rec1, r);
  double u[3];
  if constexpr (HS)
    tde_disp_point_hs(r, hrec1, X[i], Y[i], Z[i], Ss, Ds, Ts, nu, u);
  else
    tde_disp_point(r, X[i], Y[i], Z[i], Ss, Ds, Ts, nu, u);
#pragma unroll
  for (int k = 0; k < 3; ++k) out[3 * i + k] = u[k];
}

constexpr int kDsWaves = 4;  // waves per workgroup of k_disp_sum

// Displacement at nobs points summed over ns triangles, each with its own
// slip: out[3 i + c] = sum_j u_c(obs_i; triangle j, ss[j], ds[j], ts[j])
// (get_sumS, src/QDsim.py:25-53, with per-triangle slip). Lane <-> point (64
// per workgroup); wave w takes triangles w, w + 4, ..., an index uniform in
// the wave, so the records are read by scalar loads. The four partial sums
// are added in wave order, which fixes the result bit for bit.
template <bool HS>
__global__ void __launch_bounds__(64 * kDsWaves) k_disp_sum(
    const double* __restrict__ obs, int64_t nobs, const double* __restrict__ rec,
    const double* __restrict__ hrec, int64_t ns, const double* __restrict__ ss,
    const double* __restrict__ ds, const double* __restrict__ ts, double nu,
    double* __restrict__ out) {
  __shared__ double part[kDsWaves][3][64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t i = (int64_t)blockIdx.x * 64 + lane;
  const bool live = i < nobs;
  double acc[3] = {0.0, 0.0, 0.0};
  if (live) {
    const double x = obs[3 * i], y = obs[3 * i + 1], z = obs[3 * i + 2];
    for (int64_t j = w; j < ns; j += kDsWaves) {
      const double Ts = ts ? ts[j] : 0.0;
      double u[3];
      if constexpr (HS)
        tde_disp_point_hs(rec + kSrcRec * j, hrec + kHsRec * j, x, y, z, ss[j], ds[j], Ts, nu,
                          u);
      else
        tde_disp_point(rec + kSrcRec * j, x, y, z, ss[j], ds[j], Ts, nu, u);
      acc[0] += u[0];
      acc[1] += u[1];
      acc[2] += u[2];
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) part[w][c][lane] = acc[c];
  __syncthreads();
  if (w == 0 && live) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double s = part[0][c][lane];
      for (int k = 1; k < kDsWaves; ++k) s += part[k][c][lane];
      out[3 * i + c] = s;
    }
  }
}

static inline double poisson(double mu, double lam) { return 1.0 / (1.0 + lam / mu) / 2.0; }

// ---------------------------------------------------------------- host pool
// Grow-only device scratch for the host-pointer entry points.
struct HostShimPool {
  std::mutex mu;
  void* buf = nullptr;
  size_t cap = 0;
  hipStream_t stream = nullptr;
  int device = -1;
};
static HostShimPool g_pool;

static int pool_reserve(size_t bytes, void** out) {
  int dev = 0;
  PQ3D_HIP_CHECK(hipGetDevice(&dev));
  if (g_pool.device != dev) {
    // a different device: start a new buffer/stream there (the old one stays)
    g_pool.buf = nullptr;
    g_pool.cap = 0;
    g_pool.stream = nullptr;
    g_pool.device = dev;
  }
  if (!g_pool.stream) PQ3D_HIP_CHECK(hipStreamCreateWithFlags(&g_pool.stream, hipStreamNonBlocking));
  if (bytes > g_pool.cap) {
    if (g_pool.buf) PQ3D_HIP_CHECK(hipFree(g_pool.buf));
    g_pool.buf = nullptr;
    size_t cap = g_pool.cap ? g_pool.cap : (size_t)1 << 20;
    while (cap < bytes) cap *= 2;
    PQ3D_HIP_CHECK(hipMalloc(&g_pool.buf, cap));
    g_pool.cap = cap;
  }
  *out = g_pool.buf;
  return PQ3D_OK;
}

}  // namespace pq3d

using namespace pq3d;

extern "C" {

const char* pq3d_last_error(void) { return pq3d::g_err; }

int pq3d_version(void) { return 1; }

int pq3d_tde_source_records(const double* tri, int64_t n, double* rec) {
  clear_error();
  PQ3D_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return PQ3D_OK;
  PQ3D_REQUIRE(tri && rec, "null pointer");
  for (int64_t i = 0; i < n; ++i)
    tde_source_record_host(tri + 9 * i, tri + 9 * i + 3, tri + 9 * i + 6, rec + kSrcRec * i);
  return PQ3D_OK;
}

int pq3d_tde_hs_records(const double* tri, int64_t n, double* hrec) {
  clear_error();
  PQ3D_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return PQ3D_OK;
  PQ3D_REQUIRE(tri && hrec, "null pointer");
  for (int64_t i = 0; i < n; ++i)
    tde_hs_record_host(tri + 9 * i, tri + 9 * i + 3, tri + 9 * i + 6, hrec + kHsRec * i);
  return PQ3D_OK;
}

static int stress_points_launch(const double* d_X, const double* d_Y, const double* d_Z,
                                int64_t n, const double* d_rec1, const double* d_hrec1, double Ss,
                                double Ds, double Ts, double mu, double lam, double* d_stress,
                                double* d_strain, void* stream) {
  PQ3D_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return PQ3D_OK;
  PQ3D_REQUIRE(d_X && d_Y && d_Z && d_rec1 && d_stress && d_strain, "null pointer");
  PQ3D_REQUIRE(n <= (int64_t)1 << 31, "n too large (at most 2^31 receivers)");
  if (d_hrec1)
    k_stress_points<true><<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(
        d_X, d_Y, d_Z, n, d_rec1, d_hrec1, Ss, Ds, Ts, mu, lam, poisson(mu, lam), d_stress,
        d_strain);
  else
    k_stress_points<false><<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(
        d_X, d_Y, d_Z, n, d_rec1, nullptr, Ss, Ds, Ts, mu, lam, poisson(mu, lam), d_stress,
        d_strain);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

int pq3d_tde_stress_points(const double* d_X, const double* d_Y, const double* d_Z, int64_t n,
                           const double* d_rec1, double Ss, double Ds, double Ts, double mu,
                           double lam, double* d_stress, double* d_strain, void* stream) {
  clear_error();
  return stress_points_launch(d_X, d_Y, d_Z, n, d_rec1, nullptr, Ss, Ds, Ts, mu, lam, d_stress,
                              d_strain, stream);
}

int pq3d_tde_stress_points_hs(const double* d_X, const double* d_Y, const double* d_Z,
                              int64_t n, const double* d_rec1, const double* d_hrec1, double Ss,
                              double Ds, double Ts, double mu, double lam, double* d_stress,
                              double* d_strain, void* stream) {
  clear_error();
  PQ3D_REQUIRE(d_hrec1, "null half-space record");
  return stress_points_launch(d_X, d_Y, d_Z, n, d_rec1, d_hrec1, Ss, Ds, Ts, mu, lam, d_stress,
                              d_strain, stream);
}

static int stress_sources_launch(double x, double y, double z, const double* d_rec,
                                 const double* d_hrec, int64_t ns, double Ss, double Ds,
                                 double Ts, double mu, double lam, double* d_stress,
                                 double* d_strain, void* stream) {
  PQ3D_REQUIRE(ns >= 0, "n_sources must be >= 0");
  if (ns == 0) return PQ3D_OK;
  PQ3D_REQUIRE(d_rec && d_stress && d_strain, "null pointer");
  PQ3D_REQUIRE(ns <= (int64_t)1 << 31, "n_sources too large (at most 2^31)");
  if (d_hrec)
    k_stress_sources<true><<<blocks_for(ns, 256), 256, 0, as_stream(stream)>>>(
        x, y, z, d_rec, d_hrec, ns, Ss, Ds, Ts, mu, lam, poisson(mu, lam), d_stress, d_strain);
  else
    k_stress_sources<false><<<blocks_for(ns, 256), 256, 0, as_stream(stream)>>>(
        x, y, z, d_rec, nullptr, ns, Ss, Ds, Ts, mu, lam, poisson(mu, lam), d_stress, d_strain);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

int pq3d_tde_stress_sources(double x, double y, double z, const double* d_rec, int64_t ns,
                            double Ss, double Ds, double Ts, double mu, double lam,
                            double* d_stress, double* d_strain, void* stream) {
  clear_error();
  return stress_sources_launch(x, y, z, d_rec, nullptr, ns, Ss, Ds, Ts, mu, lam, d_stress,
                               d_strain, stream);
}

int pq3d_tde_stress_sources_hs(double x, double y, double z, const double* d_rec,
                               const double* d_hrec, int64_t ns, double Ss, double Ds, double Ts,
                               double mu, double lam, double* d_stress, double* d_strain,
                               void* stream) {
  clear_error();
  PQ3D_REQUIRE(d_hrec, "null half-space records");
  return stress_sources_launch(x, y, z, d_rec, d_hrec, ns, Ss, Ds, Ts, mu, lam, d_stress,
                               d_strain, stream);
}

static int traction_block_launch(const double* d_recv, const double* d_rec,
                                 const double* d_hrec, const int64_t* d_rows, int64_t row0,
                                 int64_t m,
                                 const int64_t* d_cols, int64_t col0, int64_t n, double mu,
                                 double lam, void* d_out, int64_t si, int64_t sj, int64_t plane,
                                 int64_t out_len, int f32, void* stream) {
  PQ3D_REQUIRE(m >= 0 && n >= 0, "m, n must be >= 0");
  if (m == 0 || n == 0) return PQ3D_OK;
  PQ3D_REQUIRE(d_recv && d_rec && d_out, "null pointer");
  PQ3D_REQUIRE(si >= 0 && sj >= 0 && plane >= 0, "strides must be >= 0");
  PQ3D_REQUIRE(5 * plane + (m - 1) * si + (n - 1) * sj < out_len,
               "output of %lld elements too small for a 6 x %lld x %lld block", (long long)out_len,
               (long long)m, (long long)n);
  const int64_t gy = (m + kTbRows - 1) / kTbRows;
  const int64_t gx = (n + kTbCols - 1) / kTbCols;
  PQ3D_REQUIRE(gx < (1LL << 31), "block too wide for one launch");
  const double nu = poisson(mu, lam);
  // grid.y is limited to 65535: fold extra row tiles into several launches
  const int64_t ymax = 65535;
  for (int64_t y0 = 0; y0 < gy; y0 += ymax) {
    const int64_t ny = (gy - y0 < ymax) ? gy - y0 : ymax;
    const int64_t r0 = y0 * kTbRows;
    const int64_t mm = (m - r0 < ny * kTbRows) ? m - r0 : ny * kTbRows;
    dim3 grid((unsigned)gx, (unsigned)ny);
    const int64_t* rr = d_rows ? d_rows + r0 : nullptr;
    hipStream_t st = as_stream(stream);
    if (f32) {
      float* o = (float*)d_out + r0 * si;
      if (d_hrec)
        k_traction_block<float, true><<<grid, kTbCols, 0, st>>>(
            d_recv, d_rec, d_hrec, rr, row0 + r0, mm, d_cols, col0, n, mu, lam, nu, o, si, sj,
            plane);
      else
        k_traction_block<float, false><<<grid, kTbCols, 0, st>>>(
            d_recv, d_rec, nullptr, rr, row0 + r0, mm, d_cols, col0, n, mu, lam, nu, o, si, sj,
            plane);
    } else {
      double* o = (double*)d_out + r0 * si;
      if (d_hrec)
        k_traction_block<double, true><<<grid, kTbCols, 0, st>>>(
            d_recv, d_rec, d_hrec, rr, row0 + r0, mm, d_cols, col0, n, mu, lam, nu, o, si, sj,
            plane);
      else
        k_traction_block<double, false><<<grid, kTbCols, 0, st>>>(
            d_recv, d_rec, nullptr, rr, row0 + r0, mm, d_cols, col0, n, mu, lam, nu, o, si, sj,
            plane);
    }
    PQ3D_LAUNCH_CHECK();
  }
  return PQ3D_OK;
}

int pq3d_tde_traction_block(const double* d_recv, const double* d_rec, const int64_t* d_rows,
                            int64_t row0, int64_t m, const int64_t* d_cols, int64_t col0,
                            int64_t n, double mu, double lam, double* d_out, int64_t si,
                            int64_t sj, int64_t plane, int64_t out_len, void* stream) {
  clear_error();
  return traction_block_launch(d_recv, d_rec, nullptr, d_rows, row0, m, d_cols, col0, n, mu, lam,
                               d_out, si, sj, plane, out_len, 0, stream);
}

int pq3d_tde_traction_block_f32(const double* d_recv, const double* d_rec,
                                const int64_t* d_rows, int64_t row0, int64_t m,
                                const int64_t* d_cols, int64_t col0, int64_t n, double mu,
                                double lam, float* d_out, int64_t si, int64_t sj, int64_t plane,
                                int64_t out_len, void* stream) {
  clear_error();
  return traction_block_launch(d_recv, d_rec, nullptr, d_rows, row0, m, d_cols, col0, n, mu, lam,
                               d_out, si, sj, plane, out_len, 1, stream);
}

int pq3d_tde_traction_block_hs(const double* d_recv, const double* d_rec, const double* d_hrec,
                               const int64_t* d_rows, int64_t row0, int64_t m,
                               const int64_t* d_cols, int64_t col0, int64_t n, double mu,
                               double lam, double* d_out, int64_t si, int64_t sj, int64_t plane,
                               int64_t out_len, void* stream) {
  clear_error();
  PQ3D_REQUIRE(d_hrec, "null half-space records");
  return traction_block_launch(d_recv, d_rec, d_hrec, d_rows, row0, m, d_cols, col0, n, mu, lam,
                               d_out, si, sj, plane, out_len, 0, stream);
}

int pq3d_tde_traction_block_hs_f32(const double* d_recv, const double* d_rec,
                                   const double* d_hrec, const int64_t* d_rows, int64_t row0,
                                   int64_t m, const int64_t* d_cols, int64_t col0, int64_t n,
                                   double mu, double lam, float* d_out, int64_t si, int64_t sj,
                                   int64_t plane, int64_t out_len, void* stream) {
  clear_error();
  PQ3D_REQUIRE(d_hrec, "null half-space records");
  return traction_block_launch(d_recv, d_rec, d_hrec, d_rows, row0, m, d_cols, col0, n, mu, lam,
                               d_out, si, sj, plane, out_len, 1, stream);
}

// ------------------------------------------------------------ displacement

int pq3d_tde_disp_points(const double* d_X, const double* d_Y, const double* d_Z, int64_t n,
                         const double* d_rec1, const double* d_hrec1, double Ss, double Ds,
                         double Ts, double nu, double* d_disp, void* stream) {
  clear_error();
  PQ3D_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return PQ3D_OK;
  PQ3D_REQUIRE(d_X && d_Y && d_Z && d_rec1 && d_disp, "null pointer");
  PQ3D_REQUIRE(n <= (int64_t)1 << 31, "n too large (at most 2^31 points)");
  PQ3D_REQUIRE(nu > -1.0 && nu < 0.5, "Poisson's ratio must lie in (-1, 0.5)");
  if (d_hrec1)
    k_disp_points<true><<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(
        d_X, d_Y, d_Z, n, d_rec1, d_hrec1, Ss, Ds, Ts, nu, d_disp);
  else
    k_disp_points<false><<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(
        d_X, d_Y, d_Z, n, d_rec1, nullptr, Ss, Ds, Ts, nu, d_disp);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

int pq3d_tde_disp_sum(const double* d_obs, int64_t nobs, const double* d_rec,
                      const double* d_hrec, int64_t ns, const double* d_ss, const double* d_ds,
                      const double* d_ts, double nu, double* d_out, void* stream) {
  clear_error();
  PQ3D_REQUIRE(nobs >= 0 && ns >= 0, "sizes must be >= 0");
  if (nobs == 0) return PQ3D_OK;
  PQ3D_REQUIRE(d_obs && d_out, "null pointer");
  PQ3D_REQUIRE(ns == 0 || (d_rec && d_ss && d_ds), "null source pointer");
  PQ3D_REQUIRE(nobs <= (int64_t)1 << 29, "nobs too large (at most 2^29 points)");
  PQ3D_REQUIRE(nu > -1.0 && nu < 0.5, "Poisson's ratio must lie in (-1, 0.5)");
  const unsigned int g = blocks_for(nobs, 64);
  if (d_hrec)
    k_disp_sum<true><<<g, 64 * kDsWaves, 0, as_stream(stream)>>>(d_obs, nobs, d_rec, d_hrec, ns,
                                                                 d_ss, d_ds, d_ts, nu, d_out);
  else
    k_disp_sum<false><<<g, 64 * kDsWaves, 0, as_stream(stream)>>>(d_obs, nobs, d_rec, nullptr, ns,
                                                                  d_ss, d_ds, d_ts, nu, d_out);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

// ------------------------------------------------ drop-in host entry points

static void fill_nan(double* p, size_t n) {
  const double q = std::nan("");
  for (size_t i = 0; i < n; ++i) p[i] = q;
}

static int host_stress_points(const double* X, const double* Y, const double* Z, size_t n,
                              const double* P1, const double* P2, const double* P3, double Ss,
                              double Ds, double Ts, double mu, double lam, bool hs,
                              double* stress, double* strain) {
  std::lock_guard<std::mutex> lk(g_pool.mu);
  const size_t nb = n * sizeof(double);
  const size_t total = 3 * nb + (kSrcRec + kHsRec) * sizeof(double) + 12 * nb;
  void* base = nullptr;
  int rc = pool_reserve(total, &base);
  if (rc) return rc;
  char* b = (char*)base;
  double* dX = (double*)b;
  double* dY = (double*)(b + nb);
  double* dZ = (double*)(b + 2 * nb);
  double* dR = (double*)(b + 3 * nb);
  double* dH = dR + kSrcRec;
  double* dS = dH + kHsRec;
  double* dE = dS + 6 * n;
  hipStream_t s = g_pool.stream;
  PQ3D_HIP_CHECK(hipMemcpyAsync(dX, X, nb, hipMemcpyHostToDevice, s));
  PQ3D_HIP_CHECK(hipMemcpyAsync(dY, Y, nb, hipMemcpyHostToDevice, s));
  PQ3D_HIP_CHECK(hipMemcpyAsync(dZ, Z, nb, hipMemcpyHostToDevice, s));
  double rec[kSrcRec + kHsRec];
  tde_source_record_host(P1, P2, P3, rec);
  if (hs) tde_hs_record_host(P1, P2, P3, rec + kSrcRec);
  PQ3D_HIP_CHECK(hipMemcpyAsync(dR, rec, (hs ? kSrcRec + kHsRec : kSrcRec) * sizeof(double),
                                hipMemcpyHostToDevice, s));
  rc = stress_points_launch(dX, dY, dZ, (int64_t)n, dR, hs ? dH : nullptr, Ss, Ds, Ts, mu, lam,
                            dS, dE, s);
  if (rc) return rc;
  PQ3D_HIP_CHECK(hipMemcpyAsync(stress, dS, 6 * nb, hipMemcpyDeviceToHost, s));
  PQ3D_HIP_CHECK(hipMemcpyAsync(strain, dE, 6 * nb, hipMemcpyDeviceToHost, s));
  PQ3D_HIP_CHECK(hipStreamSynchronize(s));
  return PQ3D_OK;
}

static int host_stress_sources(double x, double y, double z, const double* P1f,
                               const double* P2f, const double* P3f, size_t ns, double Ss,
                               double Ds, double Ts, double mu, double lam, bool hs,
                               double* stress, double* strain) {
  std::lock_guard<std::mutex> lk(g_pool.mu);
  const size_t nrec = (size_t)kSrcRec * ns, nh = hs ? (size_t)kHsRec * ns : 0;
  const size_t total = (nrec + nh + 12 * ns) * sizeof(double);
  void* base = nullptr;
  int rc = pool_reserve(total, &base);
  if (rc) return rc;
  double* dR = (double*)base;
  double* dH = dR + nrec;
  double* dS = dH + nh;
  double* dE = dS + 6 * ns;
  hipStream_t s = g_pool.stream;
  std::vector<double> recs(nrec + nh);
  for (size_t i = 0; i < ns; ++i) {
    tde_source_record_host(P1f + 3 * i, P2f + 3 * i, P3f + 3 * i, recs.data() + kSrcRec * i);
    if (hs)
      tde_hs_record_host(P1f + 3 * i, P2f + 3 * i, P3f + 3 * i, recs.data() + nrec + kHsRec * i);
  }
  PQ3D_HIP_CHECK(hipMemcpyAsync(dR, recs.data(), recs.size() * sizeof(double),
                                hipMemcpyHostToDevice, s));
  rc = stress_sources_launch(x, y, z, dR, hs ? dH : nullptr, (int64_t)ns, Ss, Ds, Ts, mu, lam, dS,
                             dE, s);
  if (rc) return rc;
  PQ3D_HIP_CHECK(hipMemcpyAsync(stress, dS, 6 * ns * sizeof(double), hipMemcpyDeviceToHost, s));
  PQ3D_HIP_CHECK(hipMemcpyAsync(strain, dE, 6 * ns * sizeof(double), hipMemcpyDeviceToHost, s));
  PQ3D_HIP_CHECK(hipStreamSynchronize(s));
  return PQ3D_OK;
}

// Same signature and output layout as the reference (src/TDstressFS_C.cpp:2048-2095):
// n receivers x 1 triangle, stress and strain n x 6 row-major (xx,yy,zz,xy,xz,yz);
// jud_halfspace selects the half-space solution (free surface z = 0). The
// reference returns nothing; errors here fill the outputs with NaN and leave the
// message in pq3d_last_error().
void TDstressFS_C(const double* X, const double* Y, const double* Z, size_t n, const double* P1,
                  const double* P2, const double* P3, double Ss, double Ds, double Ts, double mu,
                  double lambda_, bool jud_halfspace, double* stress, double* strain) {
  clear_error();
  if (n == 0) return;
  const int rc = host_stress_points(X, Y, Z, n, P1, P2, P3, Ss, Ds, Ts, mu, lambda_,
                                    jud_halfspace, stress, strain);
  if (rc) {
    fprintf(stderr, "pyquake3d_amd: %s\n", pq3d_last_error());
    fill_nan(stress, 6 * n);
    fill_nan(strain, 6 * n);
  }
}

// Same signature and layout as the reference (src/TDstressFS_C.cpp:2098-2143):
// one receiver x n_sources triangles given as flat n_sources x 3 arrays.
void TDstressEachSourceAtReceiver_C(double x, double y, double z, const double* P1_flat,
                                    const double* P2_flat, const double* P3_flat,
                                    size_t n_sources, double Ss, double Ds, double Ts, double mu,
                                    double lambda_, bool jud_halfspace, double* stress_out,
                                    double* strain_out) {
  clear_error();
  if (n_sources == 0) return;
  const int rc = host_stress_sources(x, y, z, P1_flat, P2_flat, P3_flat, n_sources, Ss, Ds, Ts,
                                     mu, lambda_, jud_halfspace, stress_out, strain_out);
  if (rc) {
    fprintf(stderr, "pyquake3d_amd: %s\n", pq3d_last_error());
    fill_nan(stress_out, 6 * n_sources);
    fill_nan(strain_out, 6 * n_sources);
  }
}

}  // extern "C"
