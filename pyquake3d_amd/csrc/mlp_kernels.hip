// Learned step-size proposal: a 6 -> 64 -> 64 -> 64 -> 1 MLP (tanh hidden
// layers, linear output) in FP32 with hand-written forward, backward and Adam
// kernels, plus the per-instance feature reduction over the fault cells.
//
// Parameter vector layout (P = 8833 floats):
//   W1 64x6 | b1 64 | W2 64x64 | b2 64 | W3 64x64 | b3 64 | W4 64 | b4 1
// (row-major: W[i][j] = output i, input j).
//
// Forward/backward: one workgroup of 64 threads (one wavefront) walks a chunk
// of samples; thread j owns neuron j of every hidden layer. Weights live in
// LDS. Each workgroup writes its gradient sum and loss sum to its own slot;
// k_mlp_adam sums the slots in order (deterministic) and applies Adam.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "pq3d_common.h"
#include "features_device.h"

namespace pq3d {

constexpr int kIn = 6, kH = 64;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH,
              oW3 = oB2 + kH, oB3 = oW3 + kH * kH, oW4 = oB3 + kH, oB4 = oW4 + kH,
              kP = oB4 + 1;

__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// grad != null: accumulate gradients of 0.5*(y - t)^2 summed over samples into
// grad[blockIdx.x*P ...] and the loss into loss[blockIdx.x]; out != null:
// write predictions.
__global__ void __launch_bounds__(64) k_mlp(const float* __restrict__ prm,
                                            const float* __restrict__ X,
                                            const float* __restrict__ T, int64_t nsamp,
                                            int64_t chunk, float* __restrict__ out,
                                            float* __restrict__ grad, float* __restrict__ loss) {
  __shared__ float W[kP];
  __shared__ float x[kIn], a1[kH], a2[kH], a3[kH], d3[kH], d2[kH];
  const int j = threadIdx.x;
  for (int p = j; p < kP; p += 64) W[p] = prm[p];
  __syncthreads();
  // gradient accumulators owned by thread j
  float gW1[kIn], gW2[kH], gW3[kH];
  float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, gW4 = 0.f, gb4 = 0.f, lsum = 0.f;
#pragma unroll
  for (int k = 0; k < kIn; ++k) gW1[k] = 0.f;
#pragma unroll
  for (int k = 0; k < kH; ++k) { gW2[k] = 0.f; gW3[k] = 0.f; }
  const int64_t s0 = (int64_t)blockIdx.x * chunk;
  const int64_t s1 = (s0 + chunk < nsamp) ? s0 + chunk : nsamp;
  for (int64_t s = s0; s < s1; ++s) {
    if (j < kIn) x[j] = X[s * kIn + j];
    __syncthreads();
    float z = W[oB1 + j];
#pragma unroll
    for (int k = 0; k < kIn; ++k) z += W[oW1 + j * kIn + k] * x[k];
    const float h1 = tanhf(z);
    a1[j] = h1;
    __syncthreads();
    z = W[oB2 + j];
    for (int k = 0; k < kH; ++k) z += W[oW2 + j * kH + k] * a1[k];
    const float h2 = tanhf(z);
    a2[j] = h2;
    __syncthreads();
    z = W[oB3 + j];
    for (int k = 0; k < kH; ++k) z += W[oW3 + j * kH + k] * a2[k];
    const float h3 = tanhf(z);
    a3[j] = h3;
    const float y = wave_sum(W[oW4 + j] * h3) + W[oB4];
    if (out && j == 0) out[s] = y;
    if (grad) {
      const float dy = y - T[s];
      lsum += 0.5f * dy * dy;
      gW4 += dy * h3;
      gb4 += dy;
      d3[j] = dy * W[oW4 + j] * (1.f - h3 * h3);
      __syncthreads();
      // dW3[i][j] += d3_i a2_j (thread j owns column j); d2_j = sum_i W3[i][j] d3_i
      float s2 = 0.f;
      for (int i = 0; i < kH; ++i) {
        gW3[i] += d3[i] * h2;
        s2 += W[oW3 + i * kH + j] * d3[i];
      }
      gb3 += d3[j];
      d2[j] = s2 * (1.f - h2 * h2);
      __syncthreads();
      float s1v = 0.f;
      for (int i = 0; i < kH; ++i) {
        gW2[i] += d2[i] * h1;
        s1v += W[oW2 + i * kH + j] * d2[i];
      }
      gb2 += d2[j];
      const float d1 = s1v * (1.f - h1 * h1);
#pragma unroll
      for (int k = 0; k < kIn; ++k) gW1[k] += d1 * x[k];
      gb1 += d1;
    }
    __syncthreads();
  }
  if (grad) {
    float* g = grad + (int64_t)blockIdx.x * kP;
#pragma unroll
    for (int k = 0; k < kIn; ++k) g[oW1 + j * kIn + k] = gW1[k];
    g[oB1 + j] = gb1;
    for (int i = 0; i < kH; ++i) g[oW2 + i * kH + j] = gW2[i];
    g[oB2 + j] = gb2;
    for (int i = 0; i < kH; ++i) g[oW3 + i * kH + j] = gW3[i];
    g[oB3 + j] = gb3;
    g[oW4 + j] = gW4;
    if (j == 0) {  // every lane holds the same gb4 and lsum
      g[oB4] = gb4;
      loss[blockIdx.x] = lsum;
    }
  }
}

// Sum the per-block gradient slots in order, scale by 1/nsamp, Adam step.
// state: m (P), v (P); step t (1-based) passed in.
__global__ void __launch_bounds__(256) k_mlp_adam(float* __restrict__ prm,
                                                  const float* __restrict__ grad, int64_t nblk,
                                                  float inv_n, float* __restrict__ m,
                                                  float* __restrict__ v, float lr, float b1,
                                                  float b2, float eps, int t,
                                                  float* __restrict__ gout) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kP) return;
  float g = 0.f;
  for (int64_t b = 0; b < nblk; ++b) g += grad[b * kP + p];
  g *= inv_n;
  if (gout) gout[p] = g;
  if (lr == 0.f) return;
  const float mm = b1 * m[p] + (1.f - b1) * g;
  const float vv = b2 * v[p] + (1.f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  const float mh = mm / (1.f - powf(b1, (float)t));
  const float vh = vv / (1.f - powf(b2, (float)t));
  prm[p] -= lr * mh / (sqrtf(vh) + eps);
}

// Per-instance solver-state features over the cells of each instance
// (element e belongs to instance inst[e]); slots and their encoding in
// features_device.h. Contiguous cell ranges per workgroup, LDS slots, one
// global atomic per touched instance and slot.
constexpr int kFeatMaxInst = 1024;

__global__ void __launch_bounds__(256) k_features(int64_t ne, int nb,
                                                  const int32_t* __restrict__ inst,
                                                  const double* __restrict__ slipv,
                                                  const double* __restrict__ psi,
                                                  const double* __restrict__ pb,
                                                  const double* __restrict__ pf0, double V0,
                                                  unsigned long long* __restrict__ acc) {
  __shared__ unsigned long long sh[4 * kFeatMaxInst];
  __shared__ int touched[kFeatMaxInst];
  for (int q = threadIdx.x; q < 4 * nb; q += blockDim.x) sh[q] = 0ull;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) touched[q] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  // every workgroup takes one contiguous range of cells, so (cells being
  // stored instance-major) it touches one or two instances and adds only
  // their slots to the global ones; every wavefront takes 64 consecutive
  // cells at a time and, when they all belong to one instance (the usual
  // case), reduces them in registers with one LDS atomic per slot, else adds
  // cell by cell. Maxima and integer sums are exact, so any order gives the
  // same slots.
  const int64_t chunk = (ne + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < ne ? lo + chunk : ne;
  for (int64_t e0 = lo + (threadIdx.x & ~63); e0 < hi; e0 += blockDim.x) {
    const int64_t e = e0 + lane;
    const bool valid = e < hi;
    const int b = valid ? (inst ? inst[e] : 0) : -1;
    FeatKeys k{0ull, 0ull, 0, false};
    if (valid) k = feature_keys(slipv[e], psi[e], pb[e], pf0[e], V0);
    feature_add_wave(b, k, sh, touched);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < 4 * nb; q += blockDim.x) {
    if (!touched[q >> 2]) continue;
    const int slot = q & 3;
    if (slot < 2)
      atomicMax(&acc[q], sh[q]);
    else if (sh[q])
      atomicAdd(&acc[q], sh[q]);
  }
}

__global__ void k_feat_clear(unsigned long long* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Learned-state record of the device controller (doubles per instance; the
// same layout as in the time-step kernels): learned flag of the current
// step, last accepted max(c1, c2), last dtnext / h, dtnext base, features.
enum { kLsLearned = 0, kLsRatio = 1, kLsGrowth = 2, kLsBase = 3, kLsFeat = 4, kLsW = 16 };
// Normalisation block (floats): input mean[6], input sd[6], output mean,
// output sd, margin, largest log10 correction.
enum { kNmXmu = 0, kNmXsd = 6, kNmYmu = 12, kNmYsd = 13, kNmMargin = 14, kNmHi = 15, kNmW = 16 };

// Step-size proposal at the start of a step, one wavefront per instance.
// Only instances beginning a fresh step (flag 0, no rejection yet) are
// touched: the six features (decoded from the pq3d_features slots, plus the
// last accepted error ratio and growth) and the dtnext base are stored in the
// learned-state record; with a network, the first attempt becomes
// base * 10^min(prediction - margin, hi) and the record is marked learned
// (a non-finite prediction leaves the reference's dtnext and the mark clear).
// The forward pass is k_mlp's, operation for operation.
__global__ void __launch_bounds__(64) k_learned_propose(
    int64_t n_inst, const unsigned long long* __restrict__ acc, const int64_t* __restrict__ ncell,
    const uint8_t* __restrict__ flag, const int32_t* __restrict__ nrj, double* __restrict__ hcur,
    double* __restrict__ lstate, const float* __restrict__ prm, const float* __restrict__ norm) {
  __shared__ float x[kIn], a1[kH], a2[kH];
  const int64_t b = blockIdx.x;
  const int j = threadIdx.x;
  if (b >= n_inst || flag[b] != 0 || nrj[b] != 0) return;
  double* ls = lstate + kLsW * b;
  const double nc = (double)ncell[b];
  double f[kIn];
  f[0] = __longlong_as_double((long long)acc[4 * b + 0]) - 1000.0;
  f[1] = 1000.0 - __longlong_as_double((long long)acc[4 * b + 1]);
  f[2] = (double)(long long)acc[4 * b + 2] / kFeatScale / nc;
  f[3] = (double)acc[4 * b + 3] / nc;
  f[4] = log10(fmax(ls[kLsRatio], 1e-12));
  f[5] = log10(ls[kLsGrowth]);
  const double base = hcur[b];
  __syncthreads();
  if (j == 0) {
#pragma unroll
    for (int q = 0; q < kIn; ++q) ls[kLsFeat + q] = f[q];
    ls[kLsBase] = base;
    ls[kLsLearned] = 0.0;
  }
  if (!prm) return;
  if (j < kIn) x[j] = ((float)f[j] - norm[kNmXmu + j]) / norm[kNmXsd + j];
  __syncthreads();
  float z = prm[oB1 + j];
#pragma unroll
  for (int k = 0; k < kIn; ++k) z += prm[oW1 + j * kIn + k] * x[k];
  a1[j] = tanhf(z);
  __syncthreads();
  z = prm[oB2 + j];
  for (int k = 0; k < kH; ++k) z += prm[oW2 + j * kH + k] * a1[k];
  a2[j] = tanhf(z);
  __syncthreads();
  z = prm[oB3 + j];
  for (int k = 0; k < kH; ++k) z += prm[oW3 + j * kH + k] * a2[k];
  const float h3 = tanhf(z);
  const float y = wave_sum(prm[oW4 + j] * h3) + prm[oB4];
  if (j == 0) {
    const double pred = (double)y * (double)norm[kNmYsd] + (double)norm[kNmYmu];
    if (isfinite(pred)) {
      const double corr = fmin(pred - (double)norm[kNmMargin], (double)norm[kNmHi]);
      hcur[b] = base * pow(10.0, corr);
      ls[kLsLearned] = 1.0;
    }
  }
}

}  // namespace pq3d

using namespace pq3d;

extern "C" {

int pq3d_mlp_num_params(void) { return kP; }

// Forward (and, when d_grad != NULL, backward) over nsamp samples X (nsamp x 6,
// FP32) with targets T; chunk samples per workgroup; d_grad has
// ceil(nsamp/chunk) x P floats, d_loss ceil(nsamp/chunk).
int pq3d_mlp_forward_backward(const float* d_prm, const float* d_X, const float* d_T,
                              int64_t nsamp, int64_t chunk, float* d_out, float* d_grad,
                              float* d_loss, void* stream) {
  clear_error();
  PQ3D_REQUIRE(nsamp >= 0 && chunk >= 1, "bad sizes");
  if (nsamp == 0) return PQ3D_OK;
  PQ3D_REQUIRE(d_prm && d_X, "null pointer");
  PQ3D_REQUIRE(!d_grad || (d_T && d_loss), "targets and loss slots needed for gradients");
  const int64_t nblk = (nsamp + chunk - 1) / chunk;
  PQ3D_REQUIRE(nblk < (1LL << 31), "too many blocks");
  k_mlp<<<(unsigned)nblk, 64, 0, as_stream(stream)>>>(d_prm, d_X, d_T, nsamp, chunk, d_out,
                                                       d_grad, d_loss);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

// Sum nblk gradient slots (scaled by 1/nsamp) into d_gout (optional) and, when
// lr > 0, take one Adam step on d_prm with moments d_m, d_v (step t >= 1).
int pq3d_mlp_adam(float* d_prm, const float* d_grad, int64_t nblk, int64_t nsamp, float* d_m,
                  float* d_v, float lr, float beta1, float beta2, float eps, int t, float* d_gout,
                  void* stream) {
  clear_error();
  PQ3D_REQUIRE(nblk >= 1 && nsamp >= 1 && t >= 1, "bad sizes");
  PQ3D_REQUIRE(d_prm && d_grad && d_m && d_v, "null pointer");
  k_mlp_adam<<<(kP + 255) / 256, 256, 0, as_stream(stream)>>>(
      d_prm, d_grad, nblk, 1.f / (float)nsamp, d_m, d_v, lr, beta1, beta2, eps, t, d_gout);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

// Per-instance features of an ensemble state (see k_features); cell e
// belongs to instance d_inst[e] (NULL: one instance); d_acc: 4*nb 64-bit
// integers, overwritten. Host finishing: max = bits(acc0) - 1000,
// min = 1000 - bits(acc1), mean = (int64)acc2 / 2^32 / n_cells,
// frac = acc3 / n_cells.
int pq3d_features(int64_t ne, int nb, const int32_t* d_inst, const double* d_slipv,
                  const double* d_psi, const double* d_b, const double* d_f0, double V0,
                  uint64_t* d_acc, void* stream) {
  clear_error();
  PQ3D_REQUIRE(ne >= 0 && nb >= 1 && nb <= kFeatMaxInst, "nb must be in [1, 1024]");
  PQ3D_REQUIRE(d_slipv && d_psi && d_b && d_f0 && d_acc, "null pointer");
  PQ3D_REQUIRE(nb == 1 || d_inst, "d_inst is required for nb > 1");
  hipStream_t s = as_stream(stream);
  k_feat_clear<<<blocks_for(4 * nb, 256), 256, 0, s>>>((unsigned long long*)d_acc, 4 * nb);
  if (ne > 0) {
    const int64_t want = (ne + 255) / 256;
    const unsigned grid = (unsigned)(want < 2048 ? want : 2048);
    k_features<<<grid, 256, 0, s>>>(ne, nb, d_inst, d_slipv, d_psi, d_b, d_f0, V0,
                                    (unsigned long long*)d_acc);
  }
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

// Step-size proposal of the device controller for n_inst instances (see
// k_learned_propose): d_acc from pq3d_features over the current state,
// d_ncell the cell count of every instance, d_lstate the learned-state
// records; d_prm NULL records features and base only, otherwise d_norm holds
// the 16 normalisation floats.
int pq3d_learned_propose(int64_t n_inst, const uint64_t* d_acc, const int64_t* d_ncell,
                         const uint8_t* d_flag, const int32_t* d_nrj, double* d_h,
                         double* d_lstate, const float* d_prm, const float* d_norm,
                         void* stream) {
  clear_error();
  PQ3D_REQUIRE(n_inst >= 1 && n_inst <= kFeatMaxInst, "n_inst must be in [1, 1024]");
  PQ3D_REQUIRE(d_acc && d_ncell && d_flag && d_nrj && d_h && d_lstate, "null pointer");
  PQ3D_REQUIRE(!d_prm || d_norm, "a network needs its normalisation block");
  k_learned_propose<<<(unsigned)n_inst, 64, 0, as_stream(stream)>>>(
      n_inst, (const unsigned long long*)d_acc, d_ncell, d_flag, d_nrj, d_h, d_lstate, d_prm,
      d_norm);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

}  // extern "C"
