// Quasi-dynamic traction-rate RHS and the adaptive Dormand-Prince RK45 step,
// FP64, one thread per fault cell (replaces src/QDsim.py:760-813 `derivative`,
// :1001-1134 `RungeKutte_solve_Dormand_Prince` and the state update of
// :960-993 `simu_forward`).
//
// Within one time step the reference freezes the elastic traction rates
// (K (V - Vpl), computed once per step, :906-922) and the slip rate V used by
// the aging law (:803). The seven RHS evaluations of an attempt are therefore
// purely cell-local, so a whole RK45 attempt -- all stages, the 5th- and
// 4th-order solutions and the embedded error ratios -- is ONE kernel that keeps
// every stage in registers; the two error maxima go to device memory with an
// order-independent atomic max.
//
// Cell state (MPa for tractions, as the reference stores them): sigma (Tno),
// tau1 (Tt1o), tau2 (Tt2o), psi (state). The arithmetic follows the
// reference's expressions term by term.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "pq3d_common.h"
#include "features_device.h"

namespace pq3d {

struct CellParams {
  double V0, mu, Cs;
};

// RHS of one cell (src/QDsim.py:760-813). Inputs in MPa; returns
// (dpsi/dt, dsigma/dt [MPa/s], dtau1/dt [MPa/s], dtau2/dt [MPa/s]).
__device__ inline void qd_rhs(double Tno, double Tt1o, double Tt2o, double state, double a,
                              double b, double dc, double f0, double V0, double mu, double Cs,
                              double AdotV1, double AdotV2, double dsigmadt,
                              double shear_loading, double slipv, double out[4]) {
  Tno = Tno * 1e6;
  Tt1o = Tt1o * 1e6;
  Tt2o = Tt2o * 1e6;
  const double e1p = exp(-state / a + Tt1o / (a * Tno));
  const double e1m = exp(-state / a - Tt1o / (a * Tno));
  const double e2p = exp(-state / a + Tt2o / (a * Tno));
  const double e2m = exp(-state / a - Tt2o / (a * Tno));
  const double dV1dtau = 2 * V0 / (a * Tno) * 0.5 * (e1p + e1m);
  const double dV2dtau = 2 * V0 / (a * Tno) * 0.5 * (e2p + e2m);
  const double dV1dsigma = -2 * V0 * Tt1o / (a * (Tno * Tno)) * 0.5 * (e1p + e1m);
  const double dV2dsigma = -2 * V0 * Tt2o / (a * (Tno * Tno)) * 0.5 * (e2p + e2m);
  const double dV1dstate = -2 * V0 / a * 0.5 * (e1p - e1m);
  const double dV2dstate = -2 * V0 / a * 0.5 * (e2p - e2m);
  double xs = (f0 - state) / b;
  xs = fmin(fmax(xs, -700.0), 700.0);
  const double dstatedt = b / dc * (V0 * exp(xs) - slipv);
  const double eta = mu / (2.0 * Cs);
  const double dtau1dt =
      (-AdotV1 + shear_loading - eta * (dV1dsigma * dsigmadt + dV1dstate * dstatedt)) /
      (1.0 + eta * dV1dtau);
  const double dtau2dt =
      (-AdotV2 + shear_loading - eta * (dV2dsigma * dsigmadt + dV2dstate * dstatedt)) /
      (1.0 + eta * dV2dtau);
  out[0] = dstatedt;
  out[1] = dsigmadt / 1e6;
  out[2] = dtau1dt / 1e6;
  out[3] = dtau2dt / 1e6;
}

// Instances whose error/max slots are first reduced in LDS within a block.
constexpr int kLdsInst = 64;

// Run record of one instance (doubles): end time, armed flag, event count,
// peak slip rate, event thresholds up and down, rejected attempts of the
// accepted steps so far, (unused), then event start times.
enum { kRunEnd = 0, kRunArmed = 1, kRunCount = 2, kRunPeak = 3, kRunVup = 4, kRunVdown = 5,
       kRunRejected = 6, kRunEv0 = 8, kRunW = 64, kRunEvents = kRunW - kRunEv0 };
// Learned-state record of one instance (doubles): learned flag of the
// current step, last accepted max(c1, c2), last dtnext / h, dtnext base of
// the current step, its six features.
enum { kLsLearned = 0, kLsRatio = 1, kLsGrowth = 2, kLsBase = 3, kLsFeat = 4, kLsW = 16 };
// Training sample row (floats): six features, target, weight.
constexpr int kSampleW = 8;

// Monotone encoding of |x| for atomic max (NaN sorts above +inf).
__device__ inline unsigned long long abs_bits(double x) {
  return (unsigned long long)__double_as_longlong(fabs(x));
}

// Dormand-Prince 5(4) coefficients (src/QDsim.py:1002-1032). a63 is a
// runtime argument: the reference uses -46732/5247 (src/QDsim.py:1019), the
// published tableau +46732/5247.
#define B21 (.2)
#define B31 (3. / 40)
#define B32 (9. / 40.)
#define B41 (44. / 45.)
#define B42 (-56. / 15)
#define B43 (32. / 9)
#define B51 (19372. / 6561.)
#define B52 (-25360 / 2187.)
#define B53 (64448. / 6561.)
#define B54 (-212. / 729.)
#define B61 (9017. / 3168.)
#define B62 (-355. / 33.)
#define B64 (49. / 176.)
#define B65 (-5103. / 18656.)
#define B71 (35. / 384.)
#define B73 (500. / 1113.)
#define B74 (125. / 192.)
#define B75 (-2187. / 6784.)
#define B76 (11. / 84.)
#define B81 (5179. / 57600.)
#define B83 (7571. / 16695.)
#define B84 (393. / 640.)
#define B85 (-92097. / 339200.)
#define B86 (187. / 2100.)
#define B87 (1. / 40.)

// One RK45 attempt of step h for n cells. y = (Tno, Tt1o, Tt2o, state) in,
// yn = 5th-order candidate out; err[4g..4g+3] = max over cells of
// |d psi/psi8|, |d tau1/tau1_8|, |d tau2/tau2_8|, |d sigma/sigma8| (bit patterns
// of non-negative doubles, NaN above +inf; zeroed by the entry point). Batched form: instance of cell i is
// inst[i] (or 0 when inst is null), with its own h and error slots.
__device__ inline void rk45_cell(
    int64_t i, const int32_t* __restrict__ inst, const double* __restrict__ hvec,
    const double* __restrict__ Tno, const double* __restrict__ Tt1o,
    const double* __restrict__ Tt2o, const double* __restrict__ st,
    const double* __restrict__ pa, const double* __restrict__ pb,
    const double* __restrict__ pdc, const double* __restrict__ pf0,
    const double* __restrict__ AdotV1, const double* __restrict__ AdotV2,
    const double* __restrict__ dsig, const double* __restrict__ shear_loading,
    const double* __restrict__ slipv, const double* __restrict__ cellpar, double* __restrict__ yTno,
    double* __restrict__ yTt1o, double* __restrict__ yTt2o, double* __restrict__ yst,
    unsigned long long* err, double B63, const uint8_t* __restrict__ flag) {
  const int g = inst ? inst[i] : 0;
  const double h = hvec[g];
  // a negative step size, or a nonzero flag (accepted or aborted), leaves the
  // instance out of this attempt: its candidate state and error slots stay
  if (h < 0.0 || (flag && flag[g])) return;
  const double V0 = cellpar[3 * g + 0], mu = cellpar[3 * g + 1], Cs = cellpar[3 * g + 2];
  const double a = pa[i], b = pb[i], dc = pdc[i], f0 = pf0[i];
  const double A1 = AdotV1[i], A2 = AdotV2[i], ds = dsig[i], sl = shear_loading[i], V = slipv[i];
  const double s0 = Tno[i], t10 = Tt1o[i], t20 = Tt2o[i], p0 = st[i];
  double k1[4], k2[4], k3[4], k4[4], k5[4], k6[4], k7[4];
  // order in k*: (dstate, dsigma, dtau1, dtau2)
  qd_rhs(s0, t10, t20, p0, a, b, dc, f0, V0, mu, Cs, A1, A2, ds, sl, V, k1);
  double s = s0 + h * B21 * k1[1];
  double t1 = t10 + h * B21 * k1[2];
  double t2 = t20 + h * B21 * k1[3];
  double p = p0 + h * B21 * k1[0];
  qd_rhs(s, t1, t2, p, a, b, dc, f0, V0, mu, Cs, A1, A2, ds, sl, V, k2);
  s = s0 + h * (B31 * k1[1] + B32 * k2[1]);
  t1 = t10 + h * (B31 * k1[2] + B32 * k2[2]);
  t2 = t20 + h * (B31 * k1[3] + B32 * k2[3]);
  p = p0 + h * (B31 * k1[0] + B32 * k2[0]);
  qd_rhs(s, t1, t2, p, a, b, dc, f0, V0, mu, Cs, A1, A2, ds, sl, V, k3);
  s = s0 + h * (B41 * k1[1] + B42 * k2[1] + B43 * k3[1]);
  t1 = t10 + h * (B41 * k1[2] + B42 * k2[2] + B43 * k3[2]);
  t2 = t20 + h * (B41 * k1[3] + B42 * k2[3] + B43 * k3[3]);
  p = p0 + h * (B41 * k1[0] + B42 * k2[0] + B43 * k3[0]);
  qd_rhs(s, t1, t2, p, a, b, dc, f0, V0, mu, Cs, A1, A2, ds, sl, V, k4);
  s = s0 + h * (B51 * k1[1] + B52 * k2[1] + B53 * k3[1] + B54 * k4[1]);
  t1 = t10 + h * (B51 * k1[2] + B52 * k2[2] + B53 * k3[2] + B54 * k4[2]);
  t2 = t20 + h * (B51 * k1[3] + B52 * k2[3] + B53 * k3[3] + B54 * k4[3]);
  p = p0 + h * (B51 * k1[0] + B52 * k2[0] + B53 * k3[0] + B54 * k4[0]);
  qd_rhs(s, t1, t2, p, a, b, dc, f0, V0, mu, Cs, A1, A2, ds, sl, V, k5);
  s = s0 + h * (B61 * k1[1] + B62 * k2[1] + B63 * k3[1] + B64 * k4[1] + B65 * k5[1]);
  t1 = t10 + h * (B61 * k1[2] + B62 * k2[2] + B63 * k3[2] + B64 * k4[2] + B65 * k5[2]);
  t2 = t20 + h * (B61 * k1[3] + B62 * k2[3] + B63 * k3[3] + B64 * k4[3] + B65 * k5[3]);
  p = p0 + h * (B61 * k1[0] + B62 * k2[0] + B63 * k3[0] + B64 * k4[0] + B65 * k5[0]);
  qd_rhs(s, t1, t2, p, a, b, dc, f0, V0, mu, Cs, A1, A2, ds, sl, V, k6);
  const double s7 = s0 + h * (B71 * k1[1] + B73 * k3[1] + B74 * k4[1] + B75 * k5[1] + B76 * k6[1]);
  const double t17 = t10 + h * (B71 * k1[2] + B73 * k3[2] + B74 * k4[2] + B75 * k5[2] + B76 * k6[2]);
  const double t27 = t20 + h * (B71 * k1[3] + B73 * k3[3] + B74 * k4[3] + B75 * k5[3] + B76 * k6[3]);
  const double p7 = p0 + h * (B71 * k1[0] + B73 * k3[0] + B74 * k4[0] + B75 * k5[0] + B76 * k6[0]);
  qd_rhs(s7, t17, t27, p7, a, b, dc, f0, V0, mu, Cs, A1, A2, ds, sl, V, k7);
  const double s8 = s0 + h * (B81 * k1[1] + B83 * k3[1] + B84 * k4[1] + B85 * k5[1] + B86 * k6[1] + B87 * k7[1]);
  const double t18 = t10 + h * (B81 * k1[2] + B83 * k3[2] + B84 * k4[2] + B85 * k5[2] + B86 * k6[2] + B87 * k7[2]);
  const double t28 = t20 + h * (B81 * k1[3] + B83 * k3[3] + B84 * k4[3] + B85 * k5[3] + B86 * k6[3] + B87 * k7[3]);
  const double p8 = p0 + h * (B81 * k1[0] + B83 * k3[0] + B84 * k4[0] + B85 * k5[0] + B86 * k6[0] + B87 * k7[0]);
  yTno[i] = s7;
  yTt1o[i] = t17;
  yTt2o[i] = t27;
  yst[i] = p7;
  // embedded error ratios (src/QDsim.py:1090-1110)
  const double e_state = fabs(fabs(p8 - p7) / p8);
  const double e_t1 = fabs(fabs(t18 - t17) / t18);
  const double e_t2 = fabs(fabs(t28 - t27) / t28);
  const double e_s = fabs(fabs(s8 - s7) / s8);
  // four separate maxima: the host combines them with the reference's own
  // (NaN-order-sensitive) max() of the three traction terms
  atomicMax(&err[4 * g + 0], abs_bits(e_state));
  atomicMax(&err[4 * g + 1], abs_bits(e_t1));
  atomicMax(&err[4 * g + 2], abs_bits(e_t2));
  atomicMax(&err[4 * g + 3], abs_bits(e_s));
}

__global__ void __launch_bounds__(256) k_rk45_attempt(
    int64_t n, const int32_t* __restrict__ inst, const double* __restrict__ hvec,
    const double* __restrict__ Tno, const double* __restrict__ Tt1o,
    const double* __restrict__ Tt2o, const double* __restrict__ st,
    const double* __restrict__ pa, const double* __restrict__ pb,
    const double* __restrict__ pdc, const double* __restrict__ pf0,
    const double* __restrict__ AdotV1, const double* __restrict__ AdotV2,
    const double* __restrict__ dsig, const double* __restrict__ shear_loading,
    const double* __restrict__ slipv, const double* __restrict__ cellpar, double* __restrict__ yTno,
    double* __restrict__ yTt1o, double* __restrict__ yTt2o, double* __restrict__ yst,
    unsigned long long* __restrict__ err, double B63, int64_t n_inst,
    const uint8_t* __restrict__ flag) {
  // per-block maxima first (LDS atomics), then one global atomic per
  // (instance, slot) per block: n cells would otherwise all hit 4 addresses
  __shared__ unsigned long long s_err[kLdsInst * 4];
  const bool lds = n_inst <= kLdsInst;
  if (lds)
    for (int t = threadIdx.x; t < n_inst * 4; t += blockDim.x) s_err[t] = 0ull;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rk45_cell(i, inst, hvec, Tno, Tt1o, Tt2o, st, pa, pb, pdc, pf0, AdotV1, AdotV2, dsig,
                       shear_loading, slipv, cellpar, yTno, yTt1o, yTt2o, yst,
                       lds ? s_err : err, B63, flag);
  __syncthreads();
  if (lds)
    for (int t = threadIdx.x; t < n_inst * 4; t += blockDim.x)
      if (s_err[t]) atomicMax(&err[t], s_err[t]);
}


// Accept a step: state <- candidate, slip rates from the regularized law per
// component, rake, V floor, slip (src/QDsim.py:960-993). vmax |= max V.
// With feature slots facc (the learned proposal's, 4 per instance), the
// committed cells also add the features of their new state
// (features_device.h), which is what the proposal of the instance's next
// step reads; the caller clears the accepted instances' slots first
// (k_clear_accepted). Every workgroup takes one contiguous range of cells
// (the grid is capped), so it adds to the global slots of one or two
// instances only.
constexpr unsigned kCommitMaxGrid = 2048;

__global__ void __launch_bounds__(256) k_rk45_commit(
    int64_t n, const int32_t* __restrict__ inst, const double* __restrict__ hvec,
    const uint8_t* __restrict__ accept, const double* __restrict__ pa,
    const double* __restrict__ cellpar, const double* __restrict__ yTno,
    const double* __restrict__ yTt1o, const double* __restrict__ yTt2o,
    const double* __restrict__ yst, double* __restrict__ Tno, double* __restrict__ Tt1o,
    double* __restrict__ Tt2o, double* __restrict__ st, double* __restrict__ slipv1,
    double* __restrict__ slipv2, double* __restrict__ slipv, double* __restrict__ rake,
    double* __restrict__ slip1, double* __restrict__ slip2,
    unsigned long long* __restrict__ vmax, int64_t n_inst, const double* __restrict__ pb,
    const double* __restrict__ pf0, double V0, unsigned long long* __restrict__ facc) {
  __shared__ unsigned long long s_v[kLdsInst];
  __shared__ unsigned long long s_f[4 * kLdsInst];
  __shared__ int touched[kLdsInst];
  const bool lds = n_inst <= kLdsInst;
  if (lds) {
    for (int t = threadIdx.x; t < n_inst; t += blockDim.x) {
      s_v[t] = 0ull;
      touched[t] = 0;
    }
    for (int t = threadIdx.x; t < 4 * n_inst; t += blockDim.x) s_f[t] = 0ull;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < n ? lo + chunk : n;
  // every wavefront walks 64 consecutive cells at a time (the feature
  // reduction needs all its lanes at every call)
  for (int64_t e0 = lo + (threadIdx.x & ~63); e0 < hi; e0 += blockDim.x) {
    const int64_t i = e0 + lane;
    const int g = (i < hi && inst) ? inst[i] : 0;
    const bool commit = i < hi && !(accept && accept[g] != 1);
    FeatKeys fk{0ull, 0ull, 0, false};
    if (commit) {
      const double h = hvec[g];
      const double V0g = cellpar[3 * g + 0];
      const double a = pa[i];
      const double s = yTno[i], t1 = yTt1o[i], t2 = yTt2o[i], p = yst[i];
      Tno[i] = s;
      Tt1o[i] = t1;
      Tt2o[i] = t2;
      st[i] = p;
      const double v1 = (2.0 * V0g) * exp(-p / a) * sinh(t1 / s / a);
      const double v2 = (2.0 * V0g) * exp(-p / a) * sinh(t2 / s / a);
      double v = sqrt(v1 * v1 + v2 * v2);
      slipv1[i] = v1;
      slipv2[i] = v2;
      rake[i] = atan2(t2, t1);
      if (v < 1e-30) v = 1e-30;
      slipv[i] = v;
      slip1[i] = slip1[i] + v1 * h;
      slip2[i] = slip2[i] + v2 * h;
      atomicMax(lds ? &s_v[g] : &vmax[g], abs_bits(v));
      if (facc) fk = feature_keys(v, p, pb[i], pf0[i], V0);
    }
    if (facc) feature_add_wave(commit ? g : -1, fk, lds ? s_f : facc, lds ? touched : nullptr);
  }
  __syncthreads();
  if (lds) {
    for (int t = threadIdx.x; t < n_inst; t += blockDim.x)
      if (s_v[t]) atomicMax(&vmax[t], s_v[t]);
    if (facc)
      for (int q = threadIdx.x; q < 4 * n_inst; q += blockDim.x) {
        if (!touched[q >> 2]) continue;
        if ((q & 3) < 2)
          atomicMax(&facc[q], s_f[q]);
        else if (s_f[q])
          atomicAdd(&facc[q], s_f[q]);
      }
  }
}

// Slip-rate deficit (src/QDsim.py:902-903): x1 = V1 - Vpl cos(rake0),
// x2 = V2 - Vpl sin(rake0), with the products pl1 = Vpl cos(rake0) and
// pl2 = Vpl sin(rake0) precomputed per cell (same rounding).
__global__ void __launch_bounds__(256) k_deficit(int64_t n, const double* __restrict__ slipv1,
                                                 const double* __restrict__ slipv2,
                                                 const double* __restrict__ pl1,
                                                 const double* __restrict__ pl2,
                                                 double* __restrict__ x1,
                                                 double* __restrict__ x2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x1[i] = slipv1[i] - pl1[i];
  x2[i] = slipv2[i] - pl2[i];
}

// Instance-major slip rates -> interleaved deficit (thread o = c*nb + b; the
// writes are contiguous, the nb reads of a wave fall in nb cache lines).
__global__ void __launch_bounds__(256) k_deficit_interleave(
    int64_t n, int64_t nb, const double* __restrict__ slipv1, const double* __restrict__ slipv2,
    const double* __restrict__ pl1, const double* __restrict__ pl2, double* __restrict__ x1,
    double* __restrict__ x2) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n * nb) return;
  const int64_t c = o / nb, b = o - c * nb, i = b * n + c;
  x1[o] = slipv1[i] - pl1[i];
  x2[o] = slipv2[i] - pl2[i];
}

// The same deficit written straight into the FP32 weight rows of the
// matrix-core product: w[c*nb + b] = x1 of cell c, instance b, and
// w[n*nb + c*nb + b] = x2 (the FP64 differences rounded once, as the
// product's own input conversion would round them).
__global__ void __launch_bounds__(256) k_deficit_w32(
    int64_t n, int64_t nb, const double* __restrict__ slipv1, const double* __restrict__ slipv2,
    const double* __restrict__ pl1, const double* __restrict__ pl2, float* __restrict__ w) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n * nb) return;
  const int64_t c = o / nb, b = o - c * nb, i = b * n + c;
  w[o] = (float)(slipv1[i] - pl1[i]);
  w[n * nb + o] = (float)(slipv2[i] - pl2[i]);
}

// dsigma/dt = K part + normal loading (src/QDsim.py:911-913), or the loading
// alone when the normal traction is fixed.
__global__ void __launch_bounds__(256) k_add_loading(int64_t n, double* __restrict__ dsig,
                                                     const double* __restrict__ load, int fix_tn) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dsig[i] = fix_tn ? load[i] : dsig[i] + load[i];
}

// Stand-alone RHS over n cells (the reference's `derivative`): out is 4 x n
// (dpsi/dt, dsigma/dt, dtau1/dt, dtau2/dt), the last three in MPa/s.
__global__ void __launch_bounds__(256) k_qd_rhs(
    int64_t n, const double* __restrict__ Tno, const double* __restrict__ Tt1o,
    const double* __restrict__ Tt2o, const double* __restrict__ st, const double* __restrict__ pa,
    const double* __restrict__ pb, const double* __restrict__ pdc, const double* __restrict__ pf0,
    const double* __restrict__ AdotV1, const double* __restrict__ AdotV2,
    const double* __restrict__ dsig, const double* __restrict__ shear_loading,
    const double* __restrict__ slipv, double V0, double mu, double Cs, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double k[4];
  qd_rhs(Tno[i], Tt1o[i], Tt2o[i], st[i], pa[i], pb[i], pdc[i], pf0[i], V0, mu, Cs, AdotV1[i],
         AdotV2[i], dsig[i], shear_loading[i], slipv[i], k);
  out[i] = k[0];
  out[n + i] = k[1];
  out[2 * n + i] = k[2];
  out[3 * n + i] = k[3];
}

__global__ void k_zero_u64(unsigned long long* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Python's max(a, b) and min(a, b) for floats (the reference's controller is
// Python): the first argument is kept unless the second compares greater
// (smaller), so a NaN in the second position is dropped.
__device__ inline double py_max(double a, double b) { return b > a ? b : a; }
__device__ inline double py_min(double a, double b) { return b < a ? b : a; }

// Accept/reject decision of the reference controller (src/QDsim.py:931-958)
// per instance still attempting (flag 0), from the error maxima of the last
// attempt: accept -> flag 1 and dtnext = min(hn1, hn2, 1.5 h); reject ->
// h = max(0.5 h, min(hn1, hn2)), and flag 2 with status 1 when h < 1e-15 or
// after more than 20 rejections. hn_i = 0.9 h (tol / R_i)^0.2,
// R1 = err_psi + 1e-10, R2 = max(err_tau1, err_tau2, err_sigma) + 1e-10.
// With a learned-state record (lstate != null, kLsW doubles per instance):
// an accepted step stores max(c1, c2) and dtnext / h there; a rejected
// attempt of a step that began with a learned proposal (lstate[kLsLearned])
// takes the PI fallback h (1/err)^0.14 prev^0.08 * 0.9 within [0.2 h, 0.9 h]
// (prev = the last accepted max(c1, c2), at least 1e-6) unless err is not
// finite, in which case the reference rule applies.
__global__ void k_rk45_control(int64_t n_inst, unsigned long long* __restrict__ err,
                               double tol, double* __restrict__ hcur, uint8_t* __restrict__ flag,
                               int32_t* __restrict__ nrj, double* __restrict__ dtnext,
                               int32_t* __restrict__ status, double* __restrict__ lstate) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_inst || flag[b]) return;
  const double e0 = __longlong_as_double((long long)err[4 * b + 0]);
  const double e1 = __longlong_as_double((long long)err[4 * b + 1]);
  const double e2 = __longlong_as_double((long long)err[4 * b + 2]);
  const double e3 = __longlong_as_double((long long)err[4 * b + 3]);
  // cleared for the next round's maxima (only attempting instances add to them)
#pragma unroll
  for (int q = 0; q < 4; ++q) err[4 * b + q] = 0ull;
  const double R1 = e0 + 1e-10;
  const double R2 = py_max(py_max(e1, e2), e3) + 1e-10;
  const double c1 = R1 / tol, c2 = R2 / tol;
  const double h = hcur[b];
  const double hn1 = h * 0.9 * pow(tol / R1, 0.2);
  const double hn2 = h * 0.9 * pow(tol / R2, 0.2);
  double* ls = lstate ? lstate + kLsW * b : nullptr;
  if (py_max(c1, c2) < 1.0 && !isnan(c1) && !isnan(c2)) {
    const double dn = py_min(py_min(hn1, hn2), 1.5 * h);
    dtnext[b] = dn;
    flag[b] = 1;
    if (ls) {
      ls[kLsRatio] = py_max(c1, c2);
      ls[kLsGrowth] = dn / h;
    }
    return;
  }
  const int32_t r = nrj[b] + 1;
  nrj[b] = r;
  const double e = py_max(c1, c2);
  double hn;
  if (ls && ls[kLsLearned] != 0.0 && isfinite(e)) {
    const double prev = fmax(ls[kLsRatio], 1e-6);
    const double hpi = h * 0.9 * pow(1.0 / e, 0.14) * pow(prev, 0.08);
    hn = fmin(fmax(0.2 * h, hpi), 0.9 * h);
  } else {
    hn = py_max(0.5 * h, py_min(hn1, hn2));
  }
  hcur[b] = hn;
  if (hn < 1e-15 || r > 20) {
    flag[b] = 2;
    status[b] = 1;
  }
}

// Before the commit of the accepted instances: clear their slip-rate maxima
// and (facc != null) their feature slots.
__global__ void k_clear_accepted(int64_t n_inst, const uint8_t* __restrict__ flag,
                                 unsigned long long* __restrict__ vmax,
                                 unsigned long long* __restrict__ facc) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b < n_inst && flag[b] == 1) {
    vmax[b] = 0ull;
    if (facc)
      for (int q = 0; q < 4; ++q) facc[4 * b + q] = 0ull;
  }
}

// After the commit: accepted instances advance t += h, take h = dtnext for
// their next step and start it with a clear rejection count. With a run
// record (run != null, kRunW doubles per instance): the instance finishes
// (flag 3, left out of later attempts and commits) once t >= its end time;
// an event starts when the step's max slip rate rises above the up
// threshold while armed (armed again below the down threshold), its start
// time is kept while there is room, and the peak slip rate is tracked.
// With a learned-state record and a sample buffer, every accepted step
// writes one training row (the six features of the step's start, then
// log10(h / dtnext_base)) at row nacc * n_inst + b while it fits.
__global__ void k_rk45_advance(int64_t n_inst, uint8_t* __restrict__ flag,
                               double* __restrict__ hcur, const double* __restrict__ dtnext,
                               int32_t* __restrict__ nrj, double* __restrict__ t,
                               int64_t* __restrict__ nacc,
                               const unsigned long long* __restrict__ vmax,
                               double* __restrict__ run, const double* __restrict__ lstate,
                               float* __restrict__ samples, int64_t sample_cap) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_inst || flag[b] != 1) return;
  const double h = hcur[b];
  const double tn = t[b] + h;
  t[b] = tn;
  const int64_t k = nacc[b];
  if (samples && lstate) {
    const int64_t row = k * n_inst + b;
    if (row < sample_cap) {
      const double* ls = lstate + kLsW * b;
      float* o = samples + kSampleW * row;
#pragma unroll
      for (int q = 0; q < 6; ++q) o[q] = (float)ls[kLsFeat + q];
      o[6] = (float)log10(h / ls[kLsBase]);
      o[7] = 1.0f;
    }
  }
  const int32_t rej = nrj[b];
  hcur[b] = dtnext[b];
  nrj[b] = 0;
  nacc[b] = k + 1;
  uint8_t f = 0;
  if (run) {
    double* r = run + kRunW * b;
    r[kRunRejected] += (double)rej;
    const double v = __longlong_as_double((long long)vmax[b]);
    if (v > r[kRunPeak] || isnan(v)) r[kRunPeak] = v;
    if (r[kRunArmed] != 0.0 && v > r[kRunVup]) {
      const int64_t ne = (int64_t)r[kRunCount];
      if (ne < kRunEvents) r[kRunEv0 + ne] = tn;
      r[kRunCount] = (double)(ne + 1);
      r[kRunArmed] = 0.0;
    } else if (v < r[kRunVdown]) {
      r[kRunArmed] = 1.0;
    }
    if (tn >= r[kRunEnd]) f = 3;
  }
  flag[b] = f;
}

}  // namespace pq3d

using namespace pq3d;

extern "C" {

int pq3d_rk45_attempt(int64_t n, const int32_t* d_inst, const double* d_h, const double* d_Tno,
                      const double* d_Tt1o, const double* d_Tt2o, const double* d_state,
                      const double* d_a, const double* d_b, const double* d_dc,
                      const double* d_f0, const double* d_AdotV1, const double* d_AdotV2,
                      const double* d_dsigmadt, const double* d_shear_loading,
                      const double* d_slipv, const double* d_cellpar, double* d_yTno,
                      double* d_yTt1o, double* d_yTt2o, double* d_ystate,
                      unsigned long long* d_err, double a63, int64_t n_inst, void* stream) {
  clear_error();
  PQ3D_REQUIRE(n >= 0 && n_inst >= 1, "bad sizes");
  if (n == 0) return PQ3D_OK;
  hipStream_t s = as_stream(stream);
  k_zero_u64<<<blocks_for(4 * n_inst, 256), 256, 0, s>>>(d_err, 4 * n_inst);
  k_rk45_attempt<<<blocks_for(n, 256), 256, 0, s>>>(
      n, d_inst, d_h, d_Tno, d_Tt1o, d_Tt2o, d_state, d_a, d_b, d_dc, d_f0, d_AdotV1, d_AdotV2,
      d_dsigmadt, d_shear_loading, d_slipv, d_cellpar, d_yTno, d_yTt1o, d_yTt2o, d_ystate, d_err,
      a63, n_inst, nullptr);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

int pq3d_rk45_commit(int64_t n, const int32_t* d_inst, const double* d_h,
                     const uint8_t* d_accept, const double* d_a, const double* d_cellpar,
                     const double* d_yTno, const double* d_yTt1o, const double* d_yTt2o,
                     const double* d_ystate, double* d_Tno, double* d_Tt1o, double* d_Tt2o,
                     double* d_state, double* d_slipv1, double* d_slipv2, double* d_slipv,
                     double* d_rake, double* d_slip1, double* d_slip2,
                     unsigned long long* d_vmax, int64_t n_inst, void* stream) {
  clear_error();
  PQ3D_REQUIRE(n >= 0 && n_inst >= 1, "bad sizes");
  if (n == 0) return PQ3D_OK;
  hipStream_t s = as_stream(stream);
  k_zero_u64<<<blocks_for(n_inst, 256), 256, 0, s>>>(d_vmax, n_inst);
  const unsigned cgrid = blocks_for(n, 256) < kCommitMaxGrid ? blocks_for(n, 256) : kCommitMaxGrid;
  k_rk45_commit<<<cgrid, 256, 0, s>>>(
      n, d_inst, d_h, d_accept, d_a, d_cellpar, d_yTno, d_yTt1o, d_yTt2o, d_ystate, d_Tno,
      d_Tt1o, d_Tt2o, d_state, d_slipv1, d_slipv2, d_slipv, d_rake, d_slip1, d_slip2, d_vmax,
      n_inst, nullptr, nullptr, 0.0, nullptr);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

int pq3d_deficit(int64_t n, const double* d_slipv1, const double* d_slipv2,
                 const double* d_pl1, const double* d_pl2, double* d_x1, double* d_x2,
                 void* stream) {
  clear_error();
  PQ3D_REQUIRE(n >= 0, "bad size");
  if (n == 0) return PQ3D_OK;
  k_deficit<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(n, d_slipv1, d_slipv2, d_pl1,
                                                                  d_pl2, d_x1, d_x2);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

int pq3d_deficit_interleave(int64_t n, int64_t nb, const double* d_slipv1,
                            const double* d_slipv2, const double* d_pl1, const double* d_pl2,
                            double* d_x1, double* d_x2, void* stream) {
  clear_error();
  PQ3D_REQUIRE(n >= 0 && nb >= 1, "bad sizes");
  if (n == 0) return PQ3D_OK;
  PQ3D_REQUIRE(d_slipv1 && d_slipv2 && d_pl1 && d_pl2 && d_x1 && d_x2, "null pointer");
  k_deficit_interleave<<<blocks_for(n * nb, 256), 256, 0, as_stream(stream)>>>(
      n, nb, d_slipv1, d_slipv2, d_pl1, d_pl2, d_x1, d_x2);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

int pq3d_deficit_w32(int64_t n, int64_t nb, const double* d_slipv1, const double* d_slipv2,
                     const double* d_pl1, const double* d_pl2, float* d_w, void* stream) {
  clear_error();
  PQ3D_REQUIRE(n >= 0 && nb >= 1, "bad sizes");
  if (n == 0) return PQ3D_OK;
  PQ3D_REQUIRE(d_slipv1 && d_slipv2 && d_pl1 && d_pl2 && d_w, "null pointer");
  PQ3D_REQUIRE(n * nb < (1LL << 40), "too many cells");
  k_deficit_w32<<<blocks_for(n * nb, 256), 256, 0, as_stream(stream)>>>(
      n, nb, d_slipv1, d_slipv2, d_pl1, d_pl2, d_w);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

int pq3d_qd_rhs(int64_t n, const double* d_Tno, const double* d_Tt1o, const double* d_Tt2o,
                const double* d_state, const double* d_a, const double* d_b, const double* d_dc,
                const double* d_f0, const double* d_AdotV1, const double* d_AdotV2,
                const double* d_dsigmadt, const double* d_shear_loading, const double* d_slipv,
                double V0, double mu, double Cs, double* d_out, void* stream) {
  clear_error();
  PQ3D_REQUIRE(n >= 0, "bad size");
  if (n == 0) return PQ3D_OK;
  k_qd_rhs<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(
      n, d_Tno, d_Tt1o, d_Tt2o, d_state, d_a, d_b, d_dc, d_f0, d_AdotV1, d_AdotV2, d_dsigmadt,
      d_shear_loading, d_slipv, V0, mu, Cs, d_out);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

int pq3d_add_loading(int64_t n, double* d_dsig, const double* d_load, int fix_tn,
                     void* stream) {
  clear_error();
  PQ3D_REQUIRE(n >= 0, "bad size");
  if (n == 0) return PQ3D_OK;
  k_add_loading<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(n, d_dsig, d_load, fix_tn);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

// Device-driven controller: `rounds` x (attempt of every instance with
// flag 0 at its h, then the reference's accept/reject decision), with no host
// synchronisation. The operator products must be current for the step.
// Per-instance arrays (n_inst): d_h (attempt step size), d_flag (0 attempting,
// 1 accepted, 2 aborted, 3 finished), d_nrj (rejections this step), d_dtnext,
// d_status (1 = aborted: dt < 1e-15 or more than 20 rejections). d_lstate:
// NULL, or the learned-state records (PQ3D_LS_W doubles per instance, see
// k_rk45_control).
int pq3d_rk45_rounds(int64_t n, const int32_t* d_inst, double* d_h, uint8_t* d_flag,
                     int32_t* d_nrj, double* d_dtnext, int32_t* d_status, int rounds,
                     double tol, const double* d_Tno, const double* d_Tt1o, const double* d_Tt2o,
                     const double* d_state, const double* d_a, const double* d_b,
                     const double* d_dc, const double* d_f0, const double* d_AdotV1,
                     const double* d_AdotV2, const double* d_dsigmadt,
                     const double* d_shear_loading, const double* d_slipv,
                     const double* d_cellpar, double* d_yTno, double* d_yTt1o, double* d_yTt2o,
                     double* d_ystate, unsigned long long* d_err, double a63, int64_t n_inst,
                     double* d_lstate, void* stream) {
  clear_error();
  PQ3D_REQUIRE(n >= 0 && n_inst >= 1 && rounds >= 1, "bad sizes");
  PQ3D_REQUIRE(tol > 0.0, "tol must be > 0");
  if (n == 0) return PQ3D_OK;
  hipStream_t s = as_stream(stream);
  // the control kernel clears the slots it reads, so one clear serves all rounds
  k_zero_u64<<<blocks_for(4 * n_inst, 256), 256, 0, s>>>(d_err, 4 * n_inst);
  for (int r = 0; r < rounds; ++r) {
    k_rk45_attempt<<<blocks_for(n, 256), 256, 0, s>>>(
        n, d_inst, d_h, d_Tno, d_Tt1o, d_Tt2o, d_state, d_a, d_b, d_dc, d_f0, d_AdotV1,
        d_AdotV2, d_dsigmadt, d_shear_loading, d_slipv, d_cellpar, d_yTno, d_yTt1o, d_yTt2o,
        d_ystate, d_err, a63, n_inst, d_flag);
    k_rk45_control<<<blocks_for(n_inst, 64), 64, 0, s>>>(n_inst, d_err, tol, d_h, d_flag, d_nrj,
                                                          d_dtnext, d_status, d_lstate);
  }
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

// Commit the instances with flag 1 (as pq3d_rk45_commit), then advance them:
// t += h, h = dtnext, rejection count 0, accepted-step count + 1, flag 0.
// Instances still attempting keep their state, h and count, and continue in
// the next call of pq3d_rk45_rounds (after the operator products are
// recomputed from the same, unchanged state). Optional (NULL: unused):
// d_run, per-instance run records (PQ3D_RUN_W doubles: end time, event
// detection, peak slip rate; flag 3 once the end time is reached);
// d_lstate with d_samples, one training row (PQ3D_SAMPLE_W floats) per
// accepted step at row nacc*n_inst + b, kept while row < sample_cap;
// d_facc with d_b, d_f0 and V0, the feature slots of pq3d_features (4 per
// instance): the accepted instances' slots are recomputed from their new
// state in the same pass (the other instances' slots are left as they are).
int pq3d_rk45_advance(int64_t n, const int32_t* d_inst, double* d_h, uint8_t* d_flag,
                      int32_t* d_nrj, const double* d_dtnext, double* d_t, int64_t* d_nacc,
                      const double* d_a, const double* d_cellpar, const double* d_yTno,
                      const double* d_yTt1o, const double* d_yTt2o, const double* d_ystate,
                      double* d_Tno, double* d_Tt1o, double* d_Tt2o, double* d_state,
                      double* d_slipv1, double* d_slipv2, double* d_slipv, double* d_rake,
                      double* d_slip1, double* d_slip2, unsigned long long* d_vmax,
                      double* d_run, const double* d_lstate, float* d_samples,
                      int64_t sample_cap, int64_t n_inst, const double* d_b,
                      const double* d_f0, double V0, uint64_t* d_facc_u, void* stream) {
  clear_error();
  PQ3D_REQUIRE(n >= 0 && n_inst >= 1 && sample_cap >= 0, "bad sizes");
  PQ3D_REQUIRE(!d_samples || d_lstate, "training rows need the learned-state records");
  PQ3D_REQUIRE(!d_facc_u || (d_b && d_f0 && V0 > 0.0),
               "feature slots need the b and f0 arrays and V0 > 0");
  unsigned long long* d_facc = reinterpret_cast<unsigned long long*>(d_facc_u);
  if (n == 0) return PQ3D_OK;
  hipStream_t s = as_stream(stream);
  k_clear_accepted<<<blocks_for(n_inst, 64), 64, 0, s>>>(n_inst, d_flag, d_vmax, d_facc);
  const unsigned cgrid = blocks_for(n, 256) < kCommitMaxGrid ? blocks_for(n, 256) : kCommitMaxGrid;
  k_rk45_commit<<<cgrid, 256, 0, s>>>(
      n, d_inst, d_h, d_flag, d_a, d_cellpar, d_yTno, d_yTt1o, d_yTt2o, d_ystate, d_Tno,
      d_Tt1o, d_Tt2o, d_state, d_slipv1, d_slipv2, d_slipv, d_rake, d_slip1, d_slip2, d_vmax,
      n_inst, d_b, d_f0, V0, d_facc);
  k_rk45_advance<<<blocks_for(n_inst, 64), 64, 0, s>>>(n_inst, d_flag, d_h, d_dtnext, d_nrj, d_t,
                                                        d_nacc, d_vmax, d_run, d_lstate,
                                                        d_samples, sample_cap);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

// Record widths of the device controller (run record, learned-state record,
// training row), in this order. Returns the number written.
int pq3d_controller_widths(int64_t* out, int64_t n) {
  const int64_t v[3] = {kRunW, kLsW, kSampleW};
  int64_t k = 0;
  for (; k < n && k < 3; ++k) out[k] = v[k];
  return (int)k;
}

}  // extern "C"
