// Per-cell solver-state features of the learned step-size proposal, shared
// by the feature reduction (k_features, over every cell) and the commit of
// accepted steps (k_rk45_commit, over the committed cells), so that both
// produce the same slot values bit for bit.
//
// Per instance: {max log10 V, min log10 V, sum log10(V theta/Dc),
// count(V > 1e-3)} with theta = Dc/V0 exp((psi - f0)/b) (the steady-state
// law of src/QDsim.py:642). Integer accumulation keeps the slots independent
// of the order of the atomics: slots 0/1 hold the largest bit pattern of
// (1000 + log10 V) and (1000 - log10 V) (positive doubles order like their
// bit patterns; log10 V is clamped to [-999, 999]), slot 2 the sum of
// log10(V theta/Dc) in fixed point (2^-32 units), slot 3 the count.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

namespace pq3d {

constexpr double kFeatScale = 4294967296.0;  // 2^32

struct FeatKeys {
  unsigned long long kmax, kmin;
  long long ksum;
  bool fast;
};

// The slot contributions of one cell with slip rate V and state psi.
__device__ __forceinline__ FeatKeys feature_keys(double V, double psi, double b, double f0,
                                                 double V0) {
  double lv = log10(V);
  lv = fmax(fmin(lv, 999.0), -999.0);
  // log10(V theta / Dc) = log10(V/V0) + (psi - f0)/(b ln 10)
  const double lvt = log10(V / V0) + (psi - f0) / (b * 2.302585092994046);
  FeatKeys k;
  k.kmax = (unsigned long long)__double_as_longlong(1000.0 + lv);
  k.kmin = (unsigned long long)__double_as_longlong(1000.0 - lv);
  k.ksum = llrint(lvt * kFeatScale);
  k.fast = V > 1e-3;
  return k;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}

__device__ __forceinline__ long long wave_sum_i64(long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Adds one wavefront's cells to the per-instance slots sh (4 per instance;
// LDS or global memory): b is the cell's instance, or -1 for a lane without
// a cell. When every lane holds the same instance (the usual case: cells are
// stored instance-major) the wavefront reduces in registers and lane 0 adds
// one value per slot; otherwise every lane adds its own. touched (may be
// null) marks the instances added to.
__device__ __forceinline__ void feature_add_wave(int b, const FeatKeys& k,
                                                 unsigned long long* sh, int* touched) {
  const int lane = threadIdx.x & 63;
  const int b0 = __builtin_amdgcn_readfirstlane(b);
  if (__all(b == b0)) {
    if (b0 < 0) return;
    const unsigned long long mx = wave_max_u64(k.kmax), mn = wave_max_u64(k.kmin);
    const long long sm = wave_sum_i64(k.ksum);
    const unsigned long long ct = __popcll(__ballot(k.fast));
    if (lane == 0) {
      atomicMax(&sh[4 * b0 + 0], mx);
      atomicMax(&sh[4 * b0 + 1], mn);
      atomicAdd(&sh[4 * b0 + 2], (unsigned long long)sm);
      if (ct) atomicAdd(&sh[4 * b0 + 3], ct);
      if (touched) touched[b0] = 1;
    }
  } else if (b >= 0) {
    atomicMax(&sh[4 * b + 0], k.kmax);
    atomicMax(&sh[4 * b + 1], k.kmin);
    atomicAdd(&sh[4 * b + 2], (unsigned long long)k.ksum);
    if (k.fast) atomicAdd(&sh[4 * b + 3], 1ull);
    if (touched) touched[b] = 1;
  }
}

}  // namespace pq3d
