// H-matrix kernels: batched near-field assembly, in-kernel ACA of the far
// field, and the streamed H-matvec (replaces src/Hmatrix.py:1014-1066
// ACA_worker / calc_stressfunc_* and :1445-1538 blocks_process_MVM).
//
// Index space: every vector and block range is in cluster order (the cluster
// tree's permutation applied once at upload), so every block is a pair of
// contiguous ranges [r0, r0+m) x [c0, c0+n).
//
// Operator stack. The six traction operators of a block are treated as one
// stacked matrix with 3m rows (AdotV1, AdotV2, dsigma rows) and 2n columns
// (strike-slip then dip-slip sources):
//     K[c*m + i, s*n + j] = op(c, s)[i, j],
//     op(0,0)=A1s op(1,0)=A2s op(2,0)=Bs op(0,1)=A1d op(1,1)=A2d op(2,1)=Bd,
// so that one product K [v1; v2] yields the three traction rates the time
// step needs (src/QDsim.py:906-917). Far-field blocks store K ~= U V with
// U (3m x k, column-major) and V (k x 2n, row-major); near-field blocks store
// the six m x n operators, each column-major.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <algorithm>
#include <mutex>
#include <vector>

#include "pq3d_common.h"
#include "tde_device.h"
#include "tde_hs_device.h"

namespace pq3d {

// Table row widths are explicit and exported by pq3d_layout_constants() so the
// host can check its tables against them.
// dense block table: r0, m, c0, n, base (pool offset), unused
enum { DB_R0 = 0, DB_M = 1, DB_C0 = 2, DB_N = 3, DB_BASE = 4, DB_W = 6 };
// low-rank block table: r0, m, c0, n, ubase, vbase, k, toff
enum { LB_R0 = 0, LB_M = 1, LB_C0 = 2, LB_N = 3, LB_UB = 4, LB_VB = 5, LB_K = 6, LB_TOFF = 7,
       LB_W = 8 };
// ACA workspace offsets per block: U, V, flags, unused
enum { WS_U = 0, WS_V = 1, WS_F = 2, WS_W = 4 };
// matvec work item: kind (0 dense, 1 low-rank), base, ld, plane, count, xoff
enum { IT_KIND = 0, IT_BASE = 1, IT_LD = 2, IT_PLANE = 3, IT_COUNT = 4, IT_XOFF = 5, IT_W = 6 };
// phase-1 unit (a column chunk of one or more low-rank blocks with the same
// source columns): first member row, members, first stacked column, columns,
// block n, block c0, direct flag (1: output rows are the blocks' t rows; 0:
// partial rows reduced in phase 1b), total rank (<= 64 with several members)
enum { UN_MEM = 0, UN_NMEM = 1, UN_CB = 2, UN_CC = 3, UN_N = 4, UN_C0 = 5, UN_DIRECT = 6,
       UN_KTOT = 7, UN_W = 8 };
// phase-1 unit member: V^T base, k, output row offset, first lane
enum { MB_VT = 0, MB_K = 1, MB_OUT = 2, MB_LANE = 3, MB_W = 4 };
// phase-1b row (a rank chunk of a block split over several units): t row
// offset, k, first partial row, number of units (partial rows of unit u start
// at first + u*k), first rank, ranks
enum { RD_T = 0, RD_K = 1, RD_P = 2, RD_NU = 3, RD_L0 = 4, RD_LC = 5, RD_W = 6 };
// tile: first row, rows
enum { TL_W = 2 };
// V panel packing (matrix-core product): one row per low-rank block: its V^T
// base in the pool, k, n, the panel's base, the panel's rank K, the block's
// first rank in the panel
enum { VP_SRC = 0, VP_K = 1, VP_N = 2, VP_BASE = 3, VP_KP = 4, VP_KOFF = 5, VP_W = 6 };
// panel phase-1 unit: V^T element (cb, first rank) of a 64-rank chunk of a
// panel (the chunk is 2n x kc, row-major, so the row stride is kc), ranks kc,
// output row (t row, or partial row), first stacked column cb, columns cc,
// the panel's source cluster (n, c0), direct flag
enum { PU_VB = 0, PU_KC = 1, PU_OUT = 2, PU_CB = 3, PU_CC = 4, PU_N = 5, PU_C0 = 6,
       PU_DIRECT = 7, PU_W = 8 };
constexpr int kRecvW = 12;  // receiver record: centroid + ev1, ev2, ev3

__device__ inline void load_rec_g(const double* __restrict__ src, double* r) {
#pragma unroll
  for (int k = 0; k < kSrcRec; ++k) r[k] = src[k];
}

// Unit strike and dip traction of one (receiver, source) pair, full space or
// (HS, with the source's half-space record hrec) half space.
template <bool HS>
__device__ __forceinline__ void pair_traction(const double* r, const double* __restrict__ hrec,
                                              int64_t src, const double* q, double mu,
                                              double lam, double nu, double ts[3],
                                              double td[3]) {
  if constexpr (HS)
    tde_pair_traction_sd_hs(r, hrec + (int64_t)kHsRec * src, q, q + 3, mu, lam, nu, ts, td);
  else
    tde_pair_traction_sd(r, q, q + 3, mu, lam, nu, ts, td);
}

// ------------------------------------------------------ near-field assembly
// One workgroup per dense block; lane <-> receiver row (column-major writes).
// The source records of kDaCols columns at a time are staged in LDS (row
// stride kSrcRec + 1 doubles) and read from there by the pair evaluation, so
// they take no registers: the full-space kernel runs two waves per SIMD.
constexpr int kDaCols = 64;
template <typename T, bool HS>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, 256),
                               amdgpu_waves_per_eu(HS ? 1 : 2)))
k_dense_assemble(const int64_t* __restrict__ tab, const double* __restrict__ recv,
                 const double* __restrict__ rec, const double* __restrict__ hrec, double mu,
                 double lam, double nu, T* __restrict__ pool) {
  __shared__ double srec[kDaCols * (kSrcRec + 1)];
  const int64_t* b = tab + (int64_t)blockIdx.x * DB_W;
  const int64_t r0 = b[DB_R0], m = b[DB_M], c0 = b[DB_C0], n = b[DB_N], base = b[DB_BASE];
  const int64_t mn = m * n;
  for (int64_t jc = 0; jc < n; jc += kDaCols) {
    const int64_t cw = (n - jc < kDaCols) ? n - jc : kDaCols;
    __syncthreads();  // the previous chunk's readers are done
    for (int64_t e = threadIdx.x; e < cw * kSrcRec; e += blockDim.x)
      srec[(e / kSrcRec) * (kSrcRec + 1) + e % kSrcRec] = rec[kSrcRec * (c0 + jc) + e];
    __syncthreads();
    const int64_t mc = m * cw;
    for (int64_t p = threadIdx.x; p < mc; p += blockDim.x) {
      const int64_t i = p % m, jl = p / m, j = jc + jl;
      const double* q = recv + kRecvW * (r0 + i);
      double ts[3], td[3];
      pair_traction<HS>(srec + jl * (kSrcRec + 1), hrec, c0 + j, q, mu, lam, nu, ts, td);
      T* o = pool + base + j * m + i;
      o[0 * mn] = (T)ts[0];
      o[1 * mn] = (T)ts[1];
      o[2 * mn] = (T)ts[2];
      o[3 * mn] = (T)td[0];
      o[4 * mn] = (T)td[1];
      o[5 * mn] = (T)td[2];
    }
  }
}

// ------------------------------------------------------------------- ACA
constexpr int kAcaThreads = 256;

struct ArgMax {
  double v;
  int64_t i;
};

__device__ inline ArgMax block_argmax(ArgMax a, double* sv, int64_t* si) {
  const int t = threadIdx.x;
  sv[t] = a.v;
  si[t] = a.i;
  __syncthreads();
  for (int s = kAcaThreads / 2; s > 0; s >>= 1) {
    if (t < s) {
      const double o = sv[t + s];
      const int64_t oi = si[t + s];
      // larger value wins; ties go to the smaller index (deterministic)
      if (o > sv[t] || (o == sv[t] && oi >= 0 && (si[t] < 0 || oi < si[t]))) {
        sv[t] = o;
        si[t] = oi;
      }
    }
    __syncthreads();
  }
  ArgMax r{sv[0], si[0]};
  __syncthreads();
  return r;
}

// Sum over the 64 lanes of a wavefront (every lane gets the sum; the
// butterfly order is fixed, so the result is deterministic).
__device__ inline double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline double block_sum(double v, double* sv) {
  const int t = threadIdx.x;
  sv[t] = v;
  __syncthreads();
  for (int s = kAcaThreads / 2; s > 0; s >>= 1) {
    if (t < s) sv[t] += sv[t + s];
    __syncthreads();
  }
  const double r = sv[0];
  __syncthreads();
  return r;
}

// Ordered reductions of the ACA with their loads issued eight at a time:
// each helper performs exactly the operations of the plain loop it replaces,
// in the same order (so results are bitwise those of the plain loop), but
// the loads of eight terms are in flight together instead of one after
// another.
constexpr int kAcaGroup = 8;

// acc - x[0] y[0] - x[xs] y[ys] - ... - x[(k-1) xs] y[(k-1) ys], in order of l
__device__ __forceinline__ double sub_dots_in_order(double acc, const double* __restrict__ x,
                                                    int64_t xs, const double* __restrict__ y,
                                                    int64_t ys, int k) {
  int l = 0;
  for (; l + kAcaGroup <= k; l += kAcaGroup) {
    double a[kAcaGroup], b[kAcaGroup];
#pragma unroll
    for (int u = 0; u < kAcaGroup; ++u) {
      a[u] = x[(int64_t)(l + u) * xs];
      b[u] = y[(int64_t)(l + u) * ys];
    }
#pragma unroll
    for (int u = 0; u < kAcaGroup; ++u) acc -= a[u] * b[u];
  }
  for (; l < k; ++l) acc -= x[(int64_t)l * xs] * y[(int64_t)l * ys];
  return acc;
}

// The row residual's pair of updates sharing the U row value:
// for l < k: ul = x[l xs]; a -= ul y1[l ys]; d -= ul y2[l ys]
__device__ __forceinline__ void sub_dots2_in_order(double& a, double& d,
                                                   const double* __restrict__ x, int64_t xs,
                                                   const double* __restrict__ y1,
                                                   const double* __restrict__ y2, int64_t ys,
                                                   int k) {
  int l = 0;
  for (; l + kAcaGroup <= k; l += kAcaGroup) {
    double u[kAcaGroup], p[kAcaGroup], q[kAcaGroup];
#pragma unroll
    for (int g = 0; g < kAcaGroup; ++g) {
      u[g] = x[(int64_t)(l + g) * xs];
      p[g] = y1[(int64_t)(l + g) * ys];
      q[g] = y2[(int64_t)(l + g) * ys];
    }
#pragma unroll
    for (int g = 0; g < kAcaGroup; ++g) {
      a -= u[g] * p[g];
      d -= u[g] * q[g];
    }
  }
  for (; l < k; ++l) {
    const double ul = x[(int64_t)l * xs];
    a -= ul * y1[(int64_t)l * ys];
    d -= ul * y2[(int64_t)l * ys];
  }
}

// acc + x[r0] y[r0] + x[r0 + 64] y[r0 + 64] + ... over r < R, in order of r
// (one lane's share of a wavefront dot product)
__device__ __forceinline__ double add_lane_dot_in_order(double acc, const double* __restrict__ x,
                                                        const double* __restrict__ y, int64_t r0,
                                                        int64_t R) {
  int64_t r = r0;
  for (; r + (kAcaGroup - 1) * 64 < R; r += kAcaGroup * 64) {
    double a[kAcaGroup], b[kAcaGroup];
#pragma unroll
    for (int u = 0; u < kAcaGroup; ++u) {
      a[u] = x[r + u * 64];
      b[u] = y[r + u * 64];
    }
#pragma unroll
    for (int u = 0; u < kAcaGroup; ++u) acc += a[u] * b[u];
  }
  for (; r < R; r += 64) acc += x[r] * y[r];
  return acc;
}

// Partially pivoted ACA of the stacked 3m x 2n block, one workgroup per block.
// Entries are evaluated in-kernel (TDE pair kernel), never materialised.
// Workspace: U column l at wU[ub + l*3m], V row l at wV[vb + l*2n] (FP64).
// Stops when |u_k| |v_k| <= eps |U_k V_k|_F (Frobenius estimate updated
// incrementally). The table's k column holds the block's rank cap (its
// workspace size). out_rank[b] = rank, or -1 when the cap was reached without
// convergence (the caller then stores that block densely).
// Bit masks of used rows/cols live in global scratch (flags, 1 byte each).
// The full-space kernel is held to three waves per SIMD (the pair
// evaluation's FP64 chains then overlap with other waves'; some registers
// spill to scratch: at N = 100,352 the ACA takes 0.44 s against 0.64 s at
// one wave per SIMD, 0.47 s at two, 0.46 s at four); the half-space kernel
// would spill heavily and keeps one.
template <bool HS>
__global__ void __launch_bounds__(kAcaThreads) __attribute__((amdgpu_waves_per_eu(HS ? 1 : 3))) k_aca(
    const int64_t* __restrict__ tab, const int64_t* __restrict__ wsoff,
    const double* __restrict__ recv, const double* __restrict__ rec,
    const double* __restrict__ hrec, double mu, double lam,
    double nu, double eps, int kmax, double* __restrict__ wU, double* __restrict__ wV,
    uint8_t* __restrict__ flags, int64_t* __restrict__ out_rank) {
  __shared__ double sv[kAcaThreads];
  __shared__ int64_t si[kAcaThreads];
  __shared__ double sx[kAcaThreads];
  const int tid = threadIdx.x;
  const int64_t bid = blockIdx.x;
  const int64_t* b = tab + bid * LB_W;
  const int64_t r0 = b[LB_R0], m = b[LB_M], c0 = b[LB_C0], n = b[LB_N];
  const int64_t R = 3 * m, C = 2 * n;
  // per-block rank cap (the workspace of this block holds kcap columns/rows)
  const int kcap = (int)((b[LB_K] < kmax) ? b[LB_K] : kmax);
  double* U = wU + wsoff[WS_W * bid + WS_U];
  double* V = wV + wsoff[WS_W * bid + WS_V];
  uint8_t* urow = flags + wsoff[WS_W * bid + WS_F];
  uint8_t* ucol = urow + R;
  for (int64_t r = tid; r < R; r += kAcaThreads) urow[r] = 0;
  for (int64_t c = tid; c < C; c += kAcaThreads) ucol[c] = 0;
  __syncthreads();

  double normS2 = 0.0;
  int64_t rstar = 0;  // first pivot row: component 0 of the first receiver
  int rank = 0;
  int zero_rows = 0;
  bool converged = false;
  for (int k = 0; k < kcap;) {
    // ---- residual of row rstar = (cs, is)
    const int cs = (int)(rstar / m);
    const int64_t is = rstar % m;
    double* vk = V + (int64_t)k * C;
    const double* q = recv + kRecvW * (r0 + is);
    ArgMax loc{-1.0, -1};
    for (int64_t j = tid; j < n; j += kAcaThreads) {
      double rr[kSrcRec];
      load_rec_g(rec + kSrcRec * (c0 + j), rr);
      double ts[3], td[3];
      pair_traction<HS>(rr, hrec, c0 + j, q, mu, lam, nu, ts, td);
      double a = ts[cs], d = td[cs];
      sub_dots2_in_order(a, d, U + rstar, R, V + j, V + n + j, C, k);
      vk[j] = a;
      vk[n + j] = d;
      if (!ucol[j] && fabs(a) > loc.v) { loc.v = fabs(a); loc.i = j; }
      if (!ucol[n + j] && fabs(d) > loc.v) { loc.v = fabs(d); loc.i = n + j; }
    }
    __syncthreads();
    const ArgMax piv = block_argmax(loc, sv, si);
    if (tid == 0) urow[rstar] = 1;
    __syncthreads();
    if (piv.i < 0 || !(piv.v > 0.0)) {
      // a (numerically) zero row: move to the next unused row
      ++zero_rows;
      ArgMax nx{-1.0, -1};
      for (int64_t r = tid; r < R; r += kAcaThreads)
        if (!urow[r]) { nx.v = 1.0; nx.i = r; break; }
      // smallest unused index: use -index as the value so the max picks it
      if (nx.i >= 0) nx.v = -(double)nx.i;
      else nx.v = -1e300;
      const ArgMax nr = block_argmax(nx, sv, si);
      if (nr.i < 0 || zero_rows > 3 * 64) { converged = true; break; }
      rstar = nr.i;
      continue;
    }
    const int64_t jstar = piv.i;
    const double delta = vk[jstar];
    __syncthreads();
    for (int64_t j = tid; j < C; j += kAcaThreads) vk[j] /= delta;
    if (tid == 0) ucol[jstar] = 1;
    __syncthreads();
    // ---- residual of column jstar = (ss, js): 3 stack rows per receiver
    const int ss = (int)(jstar / n);
    const int64_t js = jstar % n;
    double* uk = U + (int64_t)k * R;
    {
      double rr[kSrcRec];
      load_rec_g(rec + kSrcRec * (c0 + js), rr);
      for (int64_t i = tid; i < m; i += kAcaThreads) {
        const double* qi = recv + kRecvW * (r0 + i);
        double ts[3], td[3];
        pair_traction<HS>(rr, hrec, c0 + js, qi, mu, lam, nu, ts, td);
        for (int c = 0; c < 3; ++c) {
          const int64_t row = c * m + i;
          uk[row] = sub_dots_in_order(ss ? td[c] : ts[c], U + row, R, V + jstar, C, k);
        }
      }
    }
    __syncthreads();
    // ---- norms and Frobenius estimate
    double pu = 0.0, pv = 0.0;
    ArgMax nxt{-1.0, -1};
    for (int64_t r = tid; r < R; r += kAcaThreads) {
      const double u = uk[r];
      pu += u * u;
      if (!urow[r] && fabs(u) > nxt.v) { nxt.v = fabs(u); nxt.i = r; }
    }
    for (int64_t c = tid; c < C; c += kAcaThreads) pv += vk[c] * vk[c];
    const double nu2 = block_sum(pu, sv);
    const double nv2 = block_sum(pv, sv);
    // cross terms sum_l (U_l . u_k)(V_l . v_k) of the Frobenius estimate:
    // each wavefront forms whole dot products for l = wave, wave + 4, ...
    // (wave-level reductions, no workgroup barrier), one slot per l in LDS;
    // the slots of a chunk of kAcaThreads ranks are then added in order of l
    // by every thread. A rank step needs two barriers per chunk instead of
    // two workgroup reductions per previous rank.
    double cross = 0.0;
    const int wave = tid >> 6, lane = tid & 63;
    for (int l0 = 0; l0 < k; l0 += kAcaThreads) {
      const int lend = (k - l0 < kAcaThreads) ? k : l0 + kAcaThreads;
      for (int l = l0 + wave; l < lend; l += kAcaThreads / 64) {
        double a = add_lane_dot_in_order(0.0, U + (int64_t)l * R, uk, lane, R);
        double c2 = add_lane_dot_in_order(0.0, V + (int64_t)l * C, vk, lane, C);
        a = wave_sum_d(a);
        c2 = wave_sum_d(c2);
        if (lane == 0) sx[l - l0] = a * c2;
      }
      __syncthreads();
      for (int l = 0; l < lend - l0; ++l) cross += sx[l];
      __syncthreads();
    }
    normS2 += 2.0 * cross + nu2 * nv2;
    ++k;
    rank = k;
    const ArgMax nr = block_argmax(nxt, sv, si);
    if (sqrt(nu2 * nv2) <= eps * sqrt(fabs(normS2))) { converged = true; break; }
    if (nr.i < 0) { converged = true; break; }  // every row used: exact
    rstar = nr.i;
  }
  if (tid == 0) out_rank[bid] = converged ? rank : -1;
}

// ------------------------------------------------------------ wide ACA
// The same partially pivoted ACA as k_aca for blocks too large for one
// workgroup (at N = 100,352 the 106 blocks with m + n >= 4,096 took 0.98 s
// of a 1.27 s build on a few CUs while the rest of the GPU idled). Each rank
// step is five launches over all wide blocks at once: the row residual and
// the column residual spread over workgroups of 256 columns / receivers, the
// cross terms one wavefront per previous rank, and two one-workgroup-per-block
// steps (pivot, norms) that reduce in exactly k_aca's order, so the factors
// and ranks equal k_aca's bit for bit. Block state (int64, kAwW per block):
enum { AW_RSTAR = 0, AW_K = 1, AW_ZR = 2, AW_STATUS = 3, AW_SKIP = 4, AW_JSTAR = 5, AW_W = 8 };
// status: 0 running, 1 converged (rank = k), 2 rank cap reached (rank -1);
// skip: the pivot row of this step was numerically zero (no column step).
// Double state per block: normS2, then the kmax cross-term products.
constexpr int kAwCols = 256;

__device__ inline int aw_kcap(const int64_t* b, int kmax) {
  return (int)((b[LB_K] < kmax) ? b[LB_K] : kmax);
}

__device__ inline void aw_finish(int64_t* st, int status, int64_t* out_rank, int64_t bid,
                                 unsigned long long* done) {
  st[AW_STATUS] = status;
  out_rank[bid] = (status == 1) ? st[AW_K] : -1;
  atomicAdd(done, 1ULL);
}

__global__ void __launch_bounds__(256) k_acaw_init(const int64_t* __restrict__ tab,
                                                   const int64_t* __restrict__ wsoff, int kmax,
                                                   uint8_t* __restrict__ flags,
                                                   int64_t* __restrict__ ist,
                                                   double* __restrict__ dst, int64_t* out_rank,
                                                   unsigned long long* done) {
  const int64_t bid = blockIdx.x;
  const int64_t* b = tab + bid * LB_W;
  const int64_t R = 3 * b[LB_M], C = 2 * b[LB_N];
  uint8_t* urow = flags + wsoff[WS_W * bid + WS_F];
  for (int64_t r = threadIdx.x; r < R + C; r += blockDim.x) urow[r] = 0;
  if (threadIdx.x == 0) {
    int64_t* st = ist + AW_W * bid;
    for (int q = 0; q < AW_W; ++q) st[q] = 0;
    dst[(int64_t)(1 + kmax) * bid] = 0.0;
    if (aw_kcap(b, kmax) < 1) aw_finish(st, 2, out_rank, bid, done);
  }
}

// Row residual of every running block: lane <-> stacked column pair (j, n+j).
template <bool HS>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, kAwCols),
                               amdgpu_waves_per_eu(HS ? 1 : 2)))
k_acaw_row(const int64_t* __restrict__ tab, const int64_t* __restrict__ wsoff,
           const double* __restrict__ recv, const double* __restrict__ rec,
           const double* __restrict__ hrec, double mu, double lam, double nu,
           const double* __restrict__ wU, double* __restrict__ wV,
           const int64_t* __restrict__ ist) {
  __shared__ double recs[kAwCols * (kSrcRec + 1)];
  const int64_t bid = blockIdx.y;
  const int64_t* st = ist + AW_W * bid;
  if (st[AW_STATUS] != 0) return;
  const int64_t* b = tab + bid * LB_W;
  const int64_t r0 = b[LB_R0], m = b[LB_M], c0 = b[LB_C0], n = b[LB_N];
  const int64_t j = (int64_t)blockIdx.x * kAwCols + threadIdx.x;
  if (j >= n) return;
  const int64_t R = 3 * m, C = 2 * n;
  const int64_t rstar = st[AW_RSTAR];
  const int k = (int)st[AW_K];
  const int cs = (int)(rstar / m);
  const int64_t is = rstar % m;
  const double* U = wU + wsoff[WS_W * bid + WS_U];
  double* V = wV + wsoff[WS_W * bid + WS_V];
  double* r = recs + threadIdx.x * (kSrcRec + 1);  // this lane's slot only
  load_rec_g(rec + kSrcRec * (c0 + j), r);
  double ts[3], td[3];
  pair_traction<HS>(r, hrec, c0 + j, recv + kRecvW * (r0 + is), mu, lam, nu, ts, td);
  double a = ts[cs], d = td[cs];
  sub_dots2_in_order(a, d, U + rstar, R, V + j, V + n + j, C, k);
  double* vk = V + (int64_t)k * C;
  vk[j] = a;
  vk[n + j] = d;
}

// Pivot column of the new row (k_aca's scan order and tie rule), or the next
// unused row when the row is numerically zero; scales the row by the pivot.
__global__ void __launch_bounds__(kAcaThreads) k_acaw_pivot(
    const int64_t* __restrict__ tab, const int64_t* __restrict__ wsoff, int kmax,
    double* __restrict__ wV, uint8_t* __restrict__ flags, int64_t* __restrict__ ist,
    int64_t* __restrict__ out_rank, unsigned long long* done) {
  __shared__ double sv[kAcaThreads];
  __shared__ int64_t si[kAcaThreads];
  const int tid = threadIdx.x;
  const int64_t bid = blockIdx.x;
  int64_t* st = ist + AW_W * bid;
  if (st[AW_STATUS] != 0) return;
  const int64_t* b = tab + bid * LB_W;
  const int64_t m = b[LB_M], n = b[LB_N];
  const int64_t R = 3 * m, C = 2 * n;
  const int k = (int)st[AW_K];
  const int64_t rstar = st[AW_RSTAR];
  double* vk = wV + wsoff[WS_W * bid + WS_V] + (int64_t)k * C;
  uint8_t* urow = flags + wsoff[WS_W * bid + WS_F];
  uint8_t* ucol = urow + R;
  ArgMax loc{-1.0, -1};
  for (int64_t j = tid; j < n; j += kAcaThreads) {
    const double a = vk[j], d = vk[n + j];
    if (!ucol[j] && fabs(a) > loc.v) { loc.v = fabs(a); loc.i = j; }
    if (!ucol[n + j] && fabs(d) > loc.v) { loc.v = fabs(d); loc.i = n + j; }
  }
  __syncthreads();
  const ArgMax piv = block_argmax(loc, sv, si);
  if (tid == 0) urow[rstar] = 1;
  __syncthreads();
  if (piv.i < 0 || !(piv.v > 0.0)) {
    ArgMax nx{-1.0, -1};
    for (int64_t r = tid; r < R; r += kAcaThreads)
      if (!urow[r]) { nx.v = 1.0; nx.i = r; break; }
    if (nx.i >= 0) nx.v = -(double)nx.i;
    else nx.v = -1e300;
    const ArgMax nr = block_argmax(nx, sv, si);
    if (tid == 0) {
      const int64_t zr = st[AW_ZR] + 1;
      st[AW_ZR] = zr;
      if (nr.i < 0 || zr > 3 * 64) {
        aw_finish(st, 1, out_rank, bid, done);
      } else {
        st[AW_RSTAR] = nr.i;
        st[AW_SKIP] = 1;
      }
    }
    return;
  }
  const int64_t jstar = piv.i;
  const double delta = vk[jstar];
  __syncthreads();
  for (int64_t j = tid; j < C; j += kAcaThreads) vk[j] /= delta;
  if (tid == 0) {
    ucol[jstar] = 1;
    st[AW_JSTAR] = jstar;
  }
}

// Column residual of every running block: lane <-> receiver (3 stack rows).
template <bool HS>
__global__ void __launch_bounds__(kAwCols) k_acaw_col(
    const int64_t* __restrict__ tab, const int64_t* __restrict__ wsoff,
    const double* __restrict__ recv, const double* __restrict__ rec,
    const double* __restrict__ hrec, double mu, double lam, double nu,
    double* __restrict__ wU, const double* __restrict__ wV, const int64_t* __restrict__ ist) {
  __shared__ double rs[kSrcRec];
  const int64_t bid = blockIdx.y;
  const int64_t* st = ist + AW_W * bid;
  if (st[AW_STATUS] != 0 || st[AW_SKIP] != 0) return;
  const int64_t* b = tab + bid * LB_W;
  const int64_t r0 = b[LB_R0], m = b[LB_M], c0 = b[LB_C0], n = b[LB_N];
  if ((int64_t)blockIdx.x * kAwCols >= m) return;
  const int64_t R = 3 * m, C = 2 * n;
  const int k = (int)st[AW_K];
  const int64_t jstar = st[AW_JSTAR];
  const int ss = (int)(jstar / n);
  const int64_t js = jstar % n;
  if (threadIdx.x < kSrcRec) rs[threadIdx.x] = rec[kSrcRec * (c0 + js) + threadIdx.x];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * kAwCols + threadIdx.x;
  if (i >= m) return;
  double* U = wU + wsoff[WS_W * bid + WS_U];
  const double* V = wV + wsoff[WS_W * bid + WS_V];
  double* uk = U + (int64_t)k * R;
  double ts[3], td[3];
  pair_traction<HS>(rs, hrec, c0 + js, recv + kRecvW * (r0 + i), mu, lam, nu, ts, td);
  for (int c = 0; c < 3; ++c) {
    const int64_t row = c * m + i;
    uk[row] = sub_dots_in_order(ss ? td[c] : ts[c], U + row, R, V + jstar, C, k);
  }
}

// Cross terms (U_l . u_k)(V_l . v_k) of the Frobenius estimate: one
// wavefront per previous rank l, in k_aca's lane order and butterfly.
__global__ void __launch_bounds__(256) k_acaw_cross(const int64_t* __restrict__ tab,
                                                    const int64_t* __restrict__ wsoff, int kmax,
                                                    const double* __restrict__ wU,
                                                    const double* __restrict__ wV,
                                                    const int64_t* __restrict__ ist,
                                                    double* __restrict__ dst) {
  const int64_t bid = blockIdx.y;
  const int64_t* st = ist + AW_W * bid;
  if (st[AW_STATUS] != 0 || st[AW_SKIP] != 0) return;
  const int k = (int)st[AW_K];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int l = blockIdx.x * 4 + wave;
  if (l >= k) return;
  const int64_t* b = tab + bid * LB_W;
  const int64_t R = 3 * b[LB_M], C = 2 * b[LB_N];
  const double* U = wU + wsoff[WS_W * bid + WS_U];
  const double* V = wV + wsoff[WS_W * bid + WS_V];
  const double* uk = U + (int64_t)k * R;
  const double* vk = V + (int64_t)k * C;
  double a = add_lane_dot_in_order(0.0, U + (int64_t)l * R, uk, lane, R);
  double c2 = add_lane_dot_in_order(0.0, V + (int64_t)l * C, vk, lane, C);
  a = wave_sum_d(a);
  c2 = wave_sum_d(c2);
  if (lane == 0) dst[(int64_t)(1 + kmax) * bid + 1 + l] = a * c2;
}

// Norms, Frobenius estimate, convergence and the next pivot row.
__global__ void __launch_bounds__(kAcaThreads) k_acaw_norm(
    const int64_t* __restrict__ tab, const int64_t* __restrict__ wsoff, int kmax, double eps,
    const double* __restrict__ wU, const double* __restrict__ wV,
    const uint8_t* __restrict__ flags, int64_t* __restrict__ ist, double* __restrict__ dst,
    int64_t* __restrict__ out_rank, unsigned long long* done) {
  __shared__ double sv[kAcaThreads];
  __shared__ int64_t si[kAcaThreads];
  const int tid = threadIdx.x;
  const int64_t bid = blockIdx.x;
  int64_t* st = ist + AW_W * bid;
  if (st[AW_STATUS] != 0) return;
  if (st[AW_SKIP] != 0) {
    __syncthreads();  // every thread has read the flag before it is cleared
    if (tid == 0) st[AW_SKIP] = 0;
    return;
  }
  const int64_t* b = tab + bid * LB_W;
  const int64_t R = 3 * b[LB_M], C = 2 * b[LB_N];
  const int k = (int)st[AW_K];
  const double* uk = wU + wsoff[WS_W * bid + WS_U] + (int64_t)k * R;
  const double* vk = wV + wsoff[WS_W * bid + WS_V] + (int64_t)k * C;
  const uint8_t* urow = flags + wsoff[WS_W * bid + WS_F];
  double* ds = dst + (int64_t)(1 + kmax) * bid;
  double pu = 0.0, pv = 0.0;
  ArgMax nxt{-1.0, -1};
  for (int64_t r = tid; r < R; r += kAcaThreads) {
    const double u = uk[r];
    pu += u * u;
    if (!urow[r] && fabs(u) > nxt.v) { nxt.v = fabs(u); nxt.i = r; }
  }
  for (int64_t c = tid; c < C; c += kAcaThreads) pv += vk[c] * vk[c];
  const double nu2 = block_sum(pu, sv);
  const double nv2 = block_sum(pv, sv);
  double cross = 0.0;
  for (int l = 0; l < k; ++l) cross += ds[1 + l];
  const double normS2 = ds[0] + (2.0 * cross + nu2 * nv2);
  const ArgMax nr = block_argmax(nxt, sv, si);
  if (tid == 0) {
    ds[0] = normS2;
    st[AW_K] = k + 1;
    if (sqrt(nu2 * nv2) <= eps * sqrt(fabs(normS2)) || nr.i < 0) {
      aw_finish(st, 1, out_rank, bid, done);
    } else if (k + 1 >= aw_kcap(b, kmax)) {
      aw_finish(st, 2, out_rank, bid, done);
    } else {
      st[AW_RSTAR] = nr.i;
    }
  }
}

// Copy the converged factors into the storage pool (type T). U goes into one
// panel per leaf cluster of the block's rows: the panel of leaf (ro, ls) (row
// offset in the block, rows) starts at 3*k*ro and holds element (o*m + i, l)
// at o*k*ls + l*ls + (i - ro) (plane-major), so each row tile of the product
// reads one contiguous range and a product without the third plane skips a
// whole sub-panel. V is stored transposed (2n x k, row-major).
template <typename T>
__global__ void __launch_bounds__(256) k_aca_pack(const int64_t* __restrict__ tab,
                                                  const int64_t* __restrict__ wsoff,
                                                  const int64_t* __restrict__ tiles,
                                                  const int64_t* __restrict__ tile_of_row,
                                                  const double* __restrict__ wU,
                                                  const double* __restrict__ wV,
                                                  T* __restrict__ pool) {
  const int64_t bid = blockIdx.x;
  const int64_t* b = tab + bid * LB_W;
  const int64_t r0 = b[LB_R0], m = b[LB_M], n = b[LB_N], k = b[LB_K];
  if (k <= 0) return;
  const double* U = wU + wsoff[WS_W * bid + WS_U];
  const double* V = wV + wsoff[WS_W * bid + WS_V];
  T* pu = pool + b[LB_UB];
  T* pv = pool + b[LB_VB];
  const int64_t m3 = 3 * m, nu = m3 * k, nvv = 2 * n * k;
  // workspace U: column l contiguous, element (r, l) at l*3m + r, r = o*m + i
  for (int64_t p = threadIdx.x; p < nu; p += blockDim.x) {
    const int64_t l = p / m3, r = p - l * m3, o = r / m, i = r - o * m;
    const int64_t* tl = tiles + TL_W * tile_of_row[r0 + i];
    const int64_t ro = tl[0] - r0, ls = tl[1];
    pu[3 * k * ro + o * k * ls + l * ls + (i - ro)] = (T)U[p];
  }
  const int64_t C = 2 * n;
  for (int64_t p = threadIdx.x; p < nvv; p += blockDim.x) {
    const int64_t c = p / k, l = p - c * k;
    pv[p] = (T)V[l * C + c];
  }
}

// V panels of the matrix-core product: every low-rank block of one source
// cluster (n columns, 2n stacked) side by side as one 2n x K matrix (K = the
// sum of their ranks), stored chunk-major: ranks [64 j, 64 j + kc_j) form a
// contiguous 2n x kc_j row-major chunk at base + 2n 64 j. Element (c, l) of
// a block's V^T (pool: src + c k + l) goes to panel rank koff + l. One
// workgroup per block; reads coalesced.
__global__ void __launch_bounds__(256) k_vpanel_pack(const int64_t* __restrict__ tab,
                                                     const float* __restrict__ pool,
                                                     float* __restrict__ vp) {
  const int64_t* b = tab + (int64_t)blockIdx.x * VP_W;
  const int64_t src = b[VP_SRC], k = b[VP_K], C = 2 * b[VP_N], base = b[VP_BASE];
  const int64_t K = b[VP_KP], koff = b[VP_KOFF];
  const int64_t tot = C * k;
  for (int64_t p = threadIdx.x; p < tot; p += blockDim.x) {
    const int64_t c = p / k, l = p - c * k, r = koff + l, j = r >> 6;
    const int64_t kc = K - 64 * j < 64 ? K - 64 * j : 64;
    vp[base + C * 64 * j + c * kc + (r - 64 * j)] = pool[src + p];
  }
}

// ------------------------------------------------------------- H-matvec
// NB right-hand sides at once (an ensemble sharing the operator). Vectors are
// interleaved by instance: x[j*NB + b]. Every stored operator element is read
// once per product whatever NB is; x and t values are wave-uniform
// (scalar-cache broadcasts). Accumulation is FP64 by fused multiply-adds in an
// order that does not depend on NB, so every right-hand side of a batched
// product equals the single-vector product bit for bit.

// Index of the wavefront within its workgroup, as a wave-uniform value: table
// rows and vector offsets derived from it then go through scalar loads and
// scalar address arithmetic.
__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// MI355X hands consecutive workgroups of a launch to its 8 XCDs in turn, and
// each XCD has its own L2. Work index w = xcd_order(blockIdx.x) gives every
// XCD a contiguous range of work indices instead, so neighbouring work items
// (units reading the same source columns, row tiles of the same blocks) share
// one L2. A bijection of [0, g) for any grid size g.
__device__ __forceinline__ int64_t xcd_order(int64_t b, int64_t g) {
  constexpr int64_t X = 8;
  const int64_t q = g / X, r = g - q * X, x = b % X, i = b / X;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Column loops are unrolled so that a group's loads are all issued before its
// multiply-adds (several loads in flight per wave). Phase 1 (one accumulator
// per right-hand side) gains from deep groups: at 16 right-hand sides a group
// of 16 columns is fastest although its weights no longer fit in scalar
// registers. Phase 2 (3 NB accumulators per lane) unrolls less. Build-time
// tuning knobs (defaults: the best of sweeps on MI355X at N = 100,352).
#ifndef PQ3D_P1_U8
#define PQ3D_P1_U8 8
#endif
#ifndef PQ3D_P2_U8
#define PQ3D_P2_U8 4
#endif
#ifndef PQ3D_P1_U16
#define PQ3D_P1_U16 16
#endif
#ifndef PQ3D_P2_U16
#define PQ3D_P2_U16 2
#endif
#ifndef PQ3D_NT_LOADS
#define PQ3D_NT_LOADS 1
#endif
// phase 2 software-pipelines its column loop for NB up to this value (with
// more right-hand sides the extra registers cost more occupancy than the
// pipelining gains)
#ifndef PQ3D_P2_PF_NB
#define PQ3D_P2_PF_NB 1
#endif

// Operator loads: every stored element is used once per product, so these
// may bypass cache retention (leaving L2 to the vectors, which are re-read).
template <typename T>
__device__ __forceinline__ T ld_op(const T* p) {
#if PQ3D_NT_LOADS
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}

#ifndef PQ3D_P1_U1
#define PQ3D_P1_U1 8
#endif
template <int NB>
constexpr int p1_unroll() {
  return NB <= 2 ? PQ3D_P1_U1 : NB <= 4 ? 4 : NB <= 8 ? PQ3D_P1_U8 : PQ3D_P1_U16;
}
template <int NB>
constexpr int p2_unroll() {
  return NB <= 2 ? 8 : NB <= 4 ? 4 : NB <= 8 ? PQ3D_P2_U8 : PQ3D_P2_U16;
}
template <int NB>
constexpr bool p2_prefetch() {
  return NB <= PQ3D_P2_PF_NB;
}
// Right-hand sides per phase-2 wavefront for 8 and 16 right-hand sides: with
// fewer than NB, each item group gets NB / NW wavefronts that split the
// right-hand sides (fewer accumulator registers per wave, more waves per
// tile, the operator loads repeated per wave). Measured on MI355X at
// N = 100,352 with 16 right-hand sides: NW 16 2.02 ms, 8 2.63 ms, 4 3.70 ms,
// so the default does not split.
#ifndef PQ3D_P2_NW8
#define PQ3D_P2_NW8 8
#endif
#ifndef PQ3D_P2_NW16
#define PQ3D_P2_NW16 16
#endif
template <int NB>
constexpr int p2_halves() {
  return NB == 16 ? 16 / PQ3D_P2_NW16 : NB == 8 ? 8 / PQ3D_P2_NW8 : 1;
}

// Phase 1: t[l][b] = sum over the unit's stacked columns c of V^T[c][l] x_b[c]
// (x = v1 for c < n, v2 for c >= n); one wavefront per unit, lane <-> (member
// block, rank l); the terms of every (l, b) are added in order of c. vp is
// the lane's V^T element of the first column, k its row stride.
template <typename T, int NB>
__device__ __forceinline__ void p1_columns(const T* __restrict__ vp, int64_t k, int64_t len,
                                           const double* __restrict__ xs, double (&acc)[NB]) {
  constexpr int U = p1_unroll<NB>();
  int64_t q = 0;
  for (; q + U <= len; q += U) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = ld_op(vp + (q + u) * k);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double vu = (double)v[u];
#pragma unroll
      for (int b = 0; b < NB; ++b) acc[b] = fma(vu, xs[(q + u) * NB + b], acc[b]);
    }
  }
  for (; q < len; ++q) {
    const double vu = (double)ld_op(vp + q * k);
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[b] = fma(vu, xs[q * NB + b], acc[b]);
  }
}

// The unit's strike columns [cb, e1) read v1, its dip columns [s2, ce) v2.
template <typename T, int NB>
__device__ __forceinline__ void p1_unit_columns(const T* __restrict__ vrow, int64_t k, int64_t cb,
                                                int64_t ce, int64_t n, int64_t c0,
                                                const double* __restrict__ v1,
                                                const double* __restrict__ v2,
                                                double (&acc)[NB]) {
  const int64_t e1 = ce < n ? ce : n;
  const int64_t s2 = cb > n ? cb : n;
  if (cb < e1) p1_columns<T, NB>(vrow + cb * k, k, e1 - cb, v1 + (c0 + cb) * NB, acc);
  if (s2 < ce) p1_columns<T, NB>(vrow + s2 * k, k, ce - s2, v2 + (c0 + s2 - n) * NB, acc);
}

template <typename T, int NB>
__global__ void __launch_bounds__(256) k_hmv_phase1(const int64_t* __restrict__ units,
                                                    int64_t nunits,
                                                    const int64_t* __restrict__ members,
                                                    const T* __restrict__ pool,
                                                    const double* __restrict__ v1,
                                                    const double* __restrict__ v2,
                                                    double* __restrict__ t,
                                                    double* __restrict__ part) {
  const int64_t u = xcd_order(blockIdx.x, gridDim.x) * (blockDim.x / 64) + wave_id();
  const int lane = threadIdx.x & 63;
  if (u >= nunits) return;
  const int64_t* it = units + u * UN_W;
  const int64_t mem0 = it[UN_MEM], nmem = it[UN_NMEM], cb = it[UN_CB], cc = it[UN_CC];
  const int64_t n = it[UN_N], c0 = it[UN_C0], ktot = it[UN_KTOT];
  double* dst = it[UN_DIRECT] ? t : part;
  const int64_t ce = cb + cc;
  if (nmem == 1) {
    // one block, lanes <-> ranks in chunks of 64
    const int64_t* mb = members + mem0 * MB_W;
    const int64_t vt = mb[MB_VT], k = mb[MB_K], out = mb[MB_OUT];
    for (int64_t l0 = 0; l0 < k; l0 += 64) {
      const int64_t l = l0 + lane;
      const int64_t lc = l < k ? l : k - 1;  // idle lanes re-read row k-1, never stored
      double acc[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) acc[b] = 0.0;
      p1_unit_columns<T, NB>(pool + vt + lc, k, cb, ce, n, c0, v1, v2, acc);
      if (l < k) {
#pragma unroll
        for (int b = 0; b < NB; ++b) dst[(out + l) * NB + b] = acc[b];
      }
    }
    return;
  }
  // several blocks (total rank <= 64): lane -> the member whose lanes hold it
  int64_t vt = 0, k = 1, out = 0, lane0 = 0;
  for (int64_t q = 0; q < nmem; ++q) {
    const int64_t* mb = members + (mem0 + q) * MB_W;
    if (q == 0 || lane >= mb[MB_LANE]) {
      vt = mb[MB_VT];
      k = mb[MB_K];
      out = mb[MB_OUT];
      lane0 = mb[MB_LANE];
    }
  }
  const bool live = lane < ktot;
  const int64_t l = live ? lane - lane0 : k - 1;  // idle lanes: last member's last rank
  double acc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = 0.0;
  p1_unit_columns<T, NB>(pool + vt + l, k, cb, ce, n, c0, v1, v2, acc);
  if (live) {
#pragma unroll
    for (int b = 0; b < NB; ++b) dst[(out + l) * NB + b] = acc[b];
  }
}

// Phase 1b: blocks split over several units: t = sum of their partial rows in
// unit order (deterministic). Element e = l*NB + b of unit q sits at
// part[(p0 + q*k)*NB + e]; one wavefront per rank chunk [l0, l0 + lc) of a
// block (about 64 elements); loads are issued eight units at a time.
template <int NB, typename E = double>
__global__ void __launch_bounds__(256) k_hmv_phase1b(const int64_t* __restrict__ red,
                                                     int64_t nred, const E* __restrict__ part,
                                                     E* __restrict__ t) {
  const int64_t r = (int64_t)blockIdx.x * (blockDim.x / 64) + wave_id();
  const int lane = threadIdx.x & 63;
  if (r >= nred) return;
  const int64_t* it = red + r * RD_W;
  const int64_t to = it[RD_T], k = it[RD_K], p0 = it[RD_P], nu = it[RD_NU];
  const int64_t l0 = it[RD_L0], lc = it[RD_LC];
  const int64_t qs = k * NB;
  for (int64_t e = l0 * NB + lane; e < (l0 + lc) * NB; e += 64) {
    const E* pp = part + p0 * NB + e;
    double s = 0.0;
    int64_t q = 0;
    for (; q + 8 <= nu; q += 8) {
      E v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = pp[(q + u) * qs];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += (double)v[u];
    }
    for (; q < nu; ++q) s += (double)pp[q * qs];
    t[to * NB + e] = (E)s;
  }
}

constexpr int kTileWaves = 4;

// Phase-2 helpers: load U columns (three planes) of an item starting at
// column c, and multiply-add them into the 3*NB accumulators.
// Phase-2 operator loads are raw buffer loads: one buffer resource per item
// (its first element and byte extent), a per-lane 32-bit byte offset for the
// lane's three plane elements (off[o] = (o plane + row) sizeof(T), fixed for
// the item) and a wave-uniform byte offset per column ((c + u) ld sizeof(T),
// a scalar register). No 64-bit per-lane address arithmetic remains in the
// column loop, and a load past the item's extent would read zero.
constexpr int kBufDword3 = 0x00020000;  // raw 32-bit data, gfx9 family
constexpr int kBufAux = PQ3D_NT_LOADS ? 2 : 0;  // non-temporal

__device__ __forceinline__ __amdgpu_buffer_rsrc_t item_rsrc(const void* p, uint64_t bytes) {
  const uint32_t nrec = bytes > 0xffffffffull ? 0xffffffffu : (uint32_t)bytes;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)nrec,
                                           kBufDword3);
}
template <typename T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff);
template <>
__device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                 int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, kBufAux));
}
template <>
__device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                   int soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, kBufAux));
}

// NP row planes are read: 3 (y1, y2, y3), or 2 when the dsigma rows (the Bs
// and Bd operators) are not wanted.
template <typename T, int U, int NP>
__device__ __forceinline__ void p2_load(__amdgpu_buffer_rsrc_t r, int ldb,
                                        const uint32_t (&off)[3], int64_t c, T (&a)[U][3]) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int soff = (int)(c + u) * ldb;
#pragma unroll
    for (int o = 0; o < NP; ++o) a[u][o] = buf_load<T>(r, off[o], soff);
  }
}

// x0: the weights of the group's first column; NB per column, of which this
// wavefront uses NW (x0 already points at its first).
template <typename T, int NB, int NW, int U, int NP>
__device__ __forceinline__ void p2_madd(const T (&a)[U][3], const double* __restrict__ x0,
                                        double (&acc)[NP * NW]) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const double* x = x0 + u * NB;
#pragma unroll
    for (int o = 0; o < NP; ++o) {
      const double ao = (double)a[u][o];
#pragma unroll
      for (int b = 0; b < NW; ++b) acc[NP * b + o] = fma(ao, x[b], acc[NP * b + o]);
    }
  }
}

// Phase 2: one workgroup per row tile (a leaf cluster of <= 64 rows); lane <->
// row. Every item is three row planes (outputs y1, y2, y3) of cnt columns
// times weights w[xoff + j] (kind 0: v1, dense strike operators; kind 2: v2,
// dense dip operators; kind 1: t, low-rank U). The tile's items are dealt to
// kTileWaves item groups (item q to group q mod kTileWaves); with H > 1 each
// group has H wavefronts, each taking NB/H of the right-hand sides. Partial
// sums meet in LDS in group order, so the sum for every right-hand side is
// formed in the same order whatever NB is. y1 = AdotV1, y2 = AdotV2,
// y3 = dsigma (K part only): element (row r, rhs b) at r*NB + b, or at
// b*ybs + r when ybs > 0 (instance-major).
#ifdef PQ3D_P2_WPE
#define PQ3D_P2_ATTR __attribute__((amdgpu_waves_per_eu(PQ3D_P2_WPE)))
#else
#define PQ3D_P2_ATTR
#endif
template <typename T, int NB, int NP>
__global__ void __launch_bounds__(64 * kTileWaves * p2_halves<NB>()) PQ3D_P2_ATTR k_hmv_phase2(
    const int64_t* __restrict__ tiles, const int64_t* __restrict__ tile_ptr,
    const int64_t* __restrict__ items, const T* __restrict__ pool,
    const double* __restrict__ v1, const double* __restrict__ v2, const double* __restrict__ t,
    double* __restrict__ y1, double* __restrict__ y2, double* __restrict__ y3, int64_t ybs) {
  constexpr int U = p2_unroll<NB>();
  constexpr int H = p2_halves<NB>();
  constexpr int NW = NB / H;  // right-hand sides per wavefront
  static_assert(NW * H == NB && 64 * kTileWaves * H <= 1024, "bad split");
  static_assert(NP == 2 || NP == 3, "two or three row planes");
  __shared__ double red[NP * NB][64];
  const int64_t tile = xcd_order(blockIdx.x, gridDim.x);
  const int64_t R0 = tiles[TL_W * tile], mt = tiles[TL_W * tile + 1];
  const int w = wave_id();
  const int g = w % kTileWaves;  // item group
  const int h = w / kTileWaves;  // right-hand-side group
  const int lane = threadIdx.x & 63;
  const bool live = lane < mt;
  // lanes past the tile's last row repeat that row (in bounds, never stored),
  // which keeps the item loop free of divergent branches
  const int row = live ? lane : (int)mt - 1;
  double acc[NP * NW];
#pragma unroll
  for (int q = 0; q < NP * NW; ++q) acc[q] = 0.0;
  for (int64_t q = tile_ptr[tile] + g; q < tile_ptr[tile + 1]; q += kTileWaves) {
    const int64_t* it = items + q * IT_W;
    const int64_t kind = it[IT_KIND], base = it[IT_BASE], ld = it[IT_LD], plane = it[IT_PLANE];
    const int64_t cnt = it[IT_COUNT], xo = it[IT_XOFF];
    // the item spans (cnt - 1) ld + 2 plane + mt elements from pool + base
    // (host-checked: under 2^31 bytes)
    const __amdgpu_buffer_rsrc_t rs =
        item_rsrc(pool + base, (uint64_t)((cnt - 1) * ld + 2 * plane + mt) * sizeof(T));
    const int ldb = (int)(ld * (int64_t)sizeof(T));
    const uint32_t off[3] = {(uint32_t)(row * sizeof(T)), (uint32_t)((plane + row) * sizeof(T)),
                             (uint32_t)((2 * plane + row) * sizeof(T))};
    const double* xs = (kind == 0 ? v1 : kind == 1 ? t : v2) + xo * NB + h * NW;
    int64_t c = 0;
    if constexpr (p2_prefetch<NB>()) {
      // software pipeline: the next group's loads are in flight while the
      // current group is multiplied
      if (U <= cnt) {
        T a[U][3];
        p2_load<T, U, NP>(rs, ldb, off, c, a);
        for (;;) {
          const bool more = c + 2 * U <= cnt;
          T an[U][3];
          if (more) p2_load<T, U, NP>(rs, ldb, off, c + U, an);
          p2_madd<T, NB, NW, U, NP>(a, xs + c * NB, acc);
          c += U;
          if (!more) break;
#pragma unroll
          for (int u = 0; u < U; ++u)
#pragma unroll
            for (int o = 0; o < 3; ++o) a[u][o] = an[u][o];
        }
      }
    } else {
      for (; c + U <= cnt; c += U) {
        T a[U][3];
        p2_load<T, U, NP>(rs, ldb, off, c, a);
        p2_madd<T, NB, NW, U, NP>(a, xs + c * NB, acc);
      }
    }
    for (; c < cnt; ++c) {
      T a[1][3];
      p2_load<T, 1, NP>(rs, ldb, off, c, a);
      p2_madd<T, NB, NW, 1, NP>(a, xs + c * NB, acc);
    }
  }
  // deterministic cross-group sum: group 0 stores, groups 1.. add in order
  for (int gg = 0; gg < kTileWaves; ++gg) {
    if (g == gg) {
#pragma unroll
      for (int q = 0; q < NP * NW; ++q) {
        double* r = &red[NP * h * NW + q][lane];
        *r = (gg == 0) ? acc[q] : *r + acc[q];
      }
    }
    __syncthreads();
  }
  if (g == 0 && live) {
    const int64_t r = R0 + lane;
    const int64_t sr = ybs > 0 ? 1 : NB, sb = ybs > 0 ? ybs : 1;
#pragma unroll
    for (int bb = 0; bb < NW; ++bb) {
      const int b = h * NW + bb;
      y1[r * sr + b * sb] = red[NP * b + 0][lane];
      y2[r * sr + b * sb] = red[NP * b + 1][lane];
      if constexpr (NP == 3) y3[r * sr + b * sb] = red[NP * b + 2][lane];
    }
  }
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Phase 2 on FP32 matrix cores for 16 right-hand sides, FP32 storage,
// NP = 2 or 3 row planes, with contiguous operator loads.
//
// v_mfma_f32_16x16x4f32 computes D (16 x 16) += A (16 x 4) B (4 x 16):
//   A: lane l holds A[l & 15][l >> 4]; B: lane l holds B[l >> 4][l & 15];
//   D: lane l, register r holds D[4 (l >> 4) + r][l & 15].
// Lane l loads the four consecutive rows s .. s+3 (s = 4 (l & 15)) of column
// c + (l >> 4) of a plane with one 16-byte load, so one load instruction
// reads four whole 64-row columns. Element q of that load is the A operand
// of product q, whose 16 A rows are therefore the tile rows 4 i + q: four
// products per plane and 4-column step cover the 64 rows. A tile with
// mt < 64 rows clamps the row group start to mt - 4, so no load leaves the
// tile's rows; rows duplicated by the clamp are stored once. B is the step's
// weights (the FP32 copies of v1, v2, or the FP32 t of phase 1; one 256-byte
// load per wavefront). Each item is summed from zero on the matrix cores and then added into
// a compensated FP32 pair (hi, lo) per output value (TwoSum: the exact
// rounding error of every hi + item addition is kept in lo), so across items the
// sum carries about twice the FP32 precision; the four item groups' hi + lo
// are added in FP64 in group order. Tiles need at least 4 rows
// (host-checked).
constexpr int kMf2Waves = kTileWaves;

__device__ __forceinline__ f32x4 buf_load4(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff) {
  return __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, kBufAux));
}

template <int NP>
struct Mf2Step {
  f32x4 a[NP];
  float b;
};

template <int NP, bool MASK>
__device__ __forceinline__ void mf2_load(__amdgpu_buffer_rsrc_t rs, const uint32_t (&off)[NP],
                                         int ldb, int64_t c, int kl, int il, int64_t cnt,
                                         const float* __restrict__ xs, Mf2Step<NP>& st) {
  const bool in = !MASK || c + kl < cnt;
  const int64_t col = in ? c + kl : cnt - 1;
  // masked columns read the item's last column (in bounds) and are zeroed
  const uint32_t back = MASK ? (uint32_t)((c + kl - col) * ldb) : 0u;
  const int soff = (int)c * ldb;
#pragma unroll
  for (int o = 0; o < NP; ++o) {
    const f32x4 v = buf_load4(rs, off[o] - back, soff);
    st.a[o] = in ? v : f32x4{0.f, 0.f, 0.f, 0.f};
  }
  const float x = xs[col * 16 + il];
  st.b = in ? x : 0.0f;
}

template <int NP>
__device__ __forceinline__ void mf2_mma(const Mf2Step<NP>& st, f32x4 (&acc)[NP][4]) {
#pragma unroll
  for (int o = 0; o < NP; ++o)
#pragma unroll
    for (int q = 0; q < 4; ++q)
      acc[o][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(st.a[o][q], st.b, acc[o][q], 0, 0, 0);
}

// COMP: items added into compensated (hi, lo) FP32 pairs; otherwise every
// item group is summed in plain FP32 on the matrix cores (fewer registers,
// more waves per SIMD).
template <int NP, bool COMP>
__global__ void __launch_bounds__(64 * kMf2Waves) k_hmv_phase2_mf16(
    const int64_t* __restrict__ tiles, const int64_t* __restrict__ tile_ptr,
    const int64_t* __restrict__ items, const float* __restrict__ pool,
    const float* __restrict__ x1, const float* __restrict__ x2, const float* __restrict__ t,
    double* __restrict__ y1, double* __restrict__ y2, double* __restrict__ y3, int64_t ybs) {
  constexpr int NB = 16;
  __shared__ double red[NP][NB][64];
  const int64_t tile = xcd_order(blockIdx.x, gridDim.x);
  const int64_t R0 = tiles[TL_W * tile], mt = tiles[TL_W * tile + 1];
  const int g = wave_id();
  const int lane = threadIdx.x & 63;
  const int kl = lane >> 4, il = lane & 15;
  // the lane's row group: rows s .. s+3 of the tile (clamped into the tile)
  const int s0 = 4 * il < (int)mt - 4 ? 4 * il : (int)mt - 4;
  f32x4 hi[NP][4], lo[NP][4];
#pragma unroll
  for (int o = 0; o < NP; ++o)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      hi[o][q] = f32x4{0.f, 0.f, 0.f, 0.f};
      lo[o][q] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
  f32x4 accd[NP][4];
  f32x4 (&acc)[NP][4] = COMP ? accd : hi;
  for (int64_t qi = tile_ptr[tile] + g; qi < tile_ptr[tile + 1]; qi += kMf2Waves) {
    const int64_t* it = items + qi * IT_W;
    const int64_t kind = it[IT_KIND], base = it[IT_BASE], ld = it[IT_LD], plane = it[IT_PLANE];
    const int64_t cnt = it[IT_COUNT], xo = it[IT_XOFF];
    // the item spans (cnt - 1) ld + (NP - 1) plane + mt elements from pool + base
    // (host-checked: under 2^31 bytes)
    const __amdgpu_buffer_rsrc_t rs = item_rsrc(
        pool + base, (uint64_t)((cnt - 1) * ld + (NP - 1) * plane + mt) * sizeof(float));
    const int ldb = (int)(ld * (int64_t)sizeof(float));
    uint32_t off[NP];
#pragma unroll
    for (int o = 0; o < NP; ++o)
      off[o] = (uint32_t)((o * plane + s0) * sizeof(float)) + (uint32_t)(kl * ldb);
    const float* xs = (kind == 0 ? x1 : kind == 1 ? t : x2) + xo * NB;
    if constexpr (COMP) {
#pragma unroll
      for (int o = 0; o < NP; ++o)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[o][q] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    int64_t c = 0;
    if (cnt >= 4) {
      Mf2Step<NP> cur;
      mf2_load<NP, false>(rs, off, ldb, 0, kl, il, cnt, xs, cur);
      for (; c + 8 <= cnt; c += 4) {
        Mf2Step<NP> nxt;
        mf2_load<NP, false>(rs, off, ldb, c + 4, kl, il, cnt, xs, nxt);
        mf2_mma<NP>(cur, acc);
        cur = nxt;
      }
      mf2_mma<NP>(cur, acc);
      c += 4;
    }
    if (c < cnt) {
      Mf2Step<NP> last;
      mf2_load<NP, true>(rs, off, ldb, c, kl, il, cnt, xs, last);
      mf2_mma<NP>(last, acc);
    }
    // compensated accumulation across items (TwoSum: the exact error of
    // hi + x goes into lo, whatever their magnitudes)
    if constexpr (COMP)
#pragma unroll
    for (int o = 0; o < NP; ++o)
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float a = hi[o][q][r], x = acc[o][q][r];
          const float sum = a + x;
          const float bb = sum - a;
          const float err = (a - (sum - bb)) + (x - bb);
          hi[o][q][r] = sum;
          lo[o][q][r] += err;
        }
  }
  // deterministic cross-group sum in FP64: group 0 stores, groups 1.. add in
  // order; each row is taken from the lane group that owns it (rows the
  // clamp duplicated are skipped)
  for (int gg = 0; gg < kMf2Waves; ++gg) {
    if (g == gg) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 4 * kl + r;
        const int si = 4 * i < (int)mt - 4 ? 4 * i : (int)mt - 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = si + q;
          if (row >= 4 * i && row < mt) {
#pragma unroll
            for (int o = 0; o < NP; ++o) {
              const double v = (double)hi[o][q][r] + (double)lo[o][q][r];
              double* p = &red[o][il][row];
              *p = (gg == 0) ? v : *p + v;
            }
          }
        }
      }
    }
    __syncthreads();
  }
  double* const ys[3] = {y1, y2, y3};
  for (int e = threadIdx.x; e < NP * NB * 64; e += 64 * kMf2Waves) {
    const int o = e / (NB * 64), m = e % (NB * 64);
    // consecutive threads write consecutive addresses in either output layout
    const int row = ybs > 0 ? m % 64 : m / NB, b = ybs > 0 ? m / 64 : m % NB;
    if (row < mt)
      ys[o][ybs > 0 ? b * ybs + R0 + row : (R0 + row) * NB + b] = red[o][b][row];
  }
}

// Phase 1 of the step-stream product: t = V^T x over V panels on FP32
// matrix cores, 16 right-hand sides. A unit is a column range [cb, cb + cc)
// of one 64-rank chunk of a panel (a contiguous cc x kc row-major block);
// its kW wavefronts take its 4-column steps in turn (wavefront g: steps g,
// g + kW, ...), so a workgroup reads one contiguous window of kW steps at a
// time, each wavefront with D steps in flight. Lane l loads ranks s .. s+3
// (s = min(4 (l & 15), kc - 4)) of column 4 step + (l >> 4) with one 16-byte
// load; element q of it is the A operand of product q (rows = ranks 4 i + q),
// B the column's 16 weights (x1 rows for stacked columns < n, else x2 rows;
// the zero row for columns past the unit). The kW partial sums meet in LDS
// and are added in wavefront order in FP32; t rows (or partial rows of a
// column-split chunk) are written as FP32.
template <int D, int kW>
__global__ void __launch_bounds__(64 * kW) k_hmv_phase1_mfp(
    const int64_t* __restrict__ units, const float* __restrict__ vp, const float* __restrict__ w,
    int64_t nrows, int64_t zrow, float* __restrict__ t, float* __restrict__ part) {
  constexpr int NB = 16;
  __shared__ float sacc[kW][16][64];
  const int64_t u = xcd_order(blockIdx.x, gridDim.x);
  const int g = wave_id();
  const int lane = threadIdx.x & 63;
  const int kl = lane >> 4, il = lane & 15;
  const int64_t* it = units + u * PU_W;
  const int64_t vb = it[PU_VB], out = it[PU_OUT], cb = it[PU_CB], cc = it[PU_CC];
  const int64_t n = it[PU_N], c0 = it[PU_C0];
  const int kc = (int)it[PU_KC];
  float* dst = it[PU_DIRECT] ? t : part;
  const int64_t nsteps = (cc + 3) / 4;
  // weights: (zrow + 1) rows of 16 floats; the chunk: cc rows of kc floats
  // (both host-checked: under 2^32 bytes)
  const __amdgpu_buffer_rsrc_t wrs = item_rsrc(w, (uint64_t)(zrow + 1) * NB * sizeof(float));
  const __amdgpu_buffer_rsrc_t rs = item_rsrc(vp + vb, (uint64_t)(cc * kc) * sizeof(float));
  const int ldb = kc * (int)sizeof(float);
  const int s0 = kc >= 4 ? (4 * il < kc - 4 ? 4 * il : kc - 4) : 0;
  const uint32_t voff = (uint32_t)(s0 * sizeof(float)) + (uint32_t)(kl * ldb);
  f32x4 acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
  struct St {
    f32x4 a;
    float b;
  };
  // branch-free: a step past the unit re-reads its last step, a column past
  // the unit reads the last column; both take the zero weight row
  auto load = [&](int64_t s, St& st) {
    const bool sv = s < nsteps;
    const int64_t sc = sv ? s : nsteps - 1;
    const int64_t col = 4 * sc + kl;
    const bool in = sv && col < cc;
    const int64_t cl = col < cc ? col : cc - 1;
    const int64_t gc = cb + cl;
    const int64_t xr = !in ? zrow : gc < n ? c0 + gc : nrows + c0 + gc - n;
    st.b = __builtin_bit_cast(
        float, __builtin_amdgcn_raw_buffer_load_b32(wrs, (uint32_t)((xr * NB + il) * 4), 0, 0));
    st.a = buf_load4(rs, voff - (uint32_t)((col - cl) * ldb), (int)(4 * sc * ldb));
  };
  St buf[D];
#pragma unroll
  for (int j = 0; j < D; ++j) load(g + (int64_t)j * kW, buf[j]);
  for (int64_t s = g; s < nsteps; s += (int64_t)D * kW) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
#pragma unroll
      for (int q = 0; q < 4; ++q)
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(buf[j].a[q], buf[j].b, acc[q], 0, 0, 0);
      load(s + (int64_t)(j + D) * kW, buf[j]);
    }
  }
  // D layout: lane l, register r of product q -> rank row 4 (l >> 4) + r of
  // product q (rank s_i + q), right-hand side l & 15
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int r = 0; r < 4; ++r) sacc[g][4 * q + r][lane] = acc[q][r];
  __syncthreads();
  if (g == 0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 4 * kl + r;
      const int si = kc >= 4 ? (4 * i < kc - 4 ? 4 * i : kc - 4) : 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        float v = sacc[0][4 * q + r][lane];
#pragma unroll
        for (int ww = 1; ww < kW; ++ww) v += sacc[ww][4 * q + r][lane];
        const int rk = si + q;
        if (rk >= 4 * i && rk < kc) dst[(out + rk) * NB + il] = v;
      }
    }
  }
}

// Phase 2 on FP32 matrix cores for 16 right-hand sides as a stream of
// 4-column steps (the default matrix-core product). The host lists every
// 4-column step of every item of a row tile (steps of one item are
// consecutive, the items of a tile follow each other), and splits a tile's
// steps into kW contiguous ranges, one per wavefront (wptr). Step record (4
// x int32): pool element of the step's first column at the tile's first row
// (plane 0), column stride, plane stride, (columns - 1) << 28 | weight row.
// Weight rows are rows of one FP32 array w: x1 rows (n), x2 rows (n), t rows
// (nt), then one row of zeros (zrow); a row holds the 16 right-hand sides.
// Every wavefront keeps D steps in flight across item boundaries: a step's
// loads are issued D steps before its products, its record 2 D steps
// before. Loads are branch-free: a step past the range re-reads the range's
// last step with the zero weight row, a column past the step's count reads
// the step's last column with the zero weight row, so both add exact zeros.
// The lane layout of the operator loads and of the products is that of
// k_hmv_phase2_mf16 (lane l loads rows s .. s+3, s = min(4 (l & 15), mt - 4),
// of column l >> 4; element q feeds product q). Each wavefront sums its steps
// in FP32; the kW partial sums meet in LDS in FP64 in wave order.
constexpr uint32_t kStepCntShift = 28;
constexpr uint32_t kStepRowMask = (1u << kStepCntShift) - 1u;

template <int NP>
struct MfsStep {
  f32x4 a[NP];
  float b;
};

template <int NP, int D, int kW>
__global__ void __launch_bounds__(64 * kW) k_hmv_phase2_mfs(
    const int4* __restrict__ steps, const int64_t* __restrict__ wptr,
    const int64_t* __restrict__ tiles, const float* __restrict__ pool,
    const float* __restrict__ w, int64_t zrow, double* __restrict__ y1,
    double* __restrict__ y2, double* __restrict__ y3, int64_t ybs) {
  constexpr int NB = 16;
  __shared__ double red[NP][NB][64];
  const int64_t tile = xcd_order(blockIdx.x, gridDim.x);
  const int64_t R0 = tiles[TL_W * tile], mt = tiles[TL_W * tile + 1];
  const int g = wave_id();
  const int lane = threadIdx.x & 63;
  const int kl = lane >> 4, il = lane & 15;
  const int s0 = 4 * il < (int)mt - 4 ? 4 * il : (int)mt - 4;
  const int64_t sb = wptr[tile * kW + g], se = wptr[tile * kW + g + 1];
  // the weight array spans (zrow + 1) rows of 16 floats (host-checked:
  // under 2^32 bytes)
  const __amdgpu_buffer_rsrc_t wrs = item_rsrc(w, (uint64_t)(zrow + 1) * NB * sizeof(float));
  f32x4 acc[NP][4];
#pragma unroll
  for (int o = 0; o < NP; ++o)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[o][q] = f32x4{0.f, 0.f, 0.f, 0.f};
  if (sb < se) {
    auto record = [&](int64_t s) -> int4 { return steps[s < se ? s : se - 1]; };
    auto load = [&](const int4 d, bool valid, MfsStep<NP>& st) {
      const uint32_t ws = (uint32_t)d.w;
      const int cnt = (int)(ws >> kStepCntShift) + 1;
      const int64_t ld = d.y, plane = d.z;
      const int c = kl < cnt ? kl : cnt - 1;
      // the step spans (cnt - 1) ld + (NP - 1) plane + mt elements (host-checked:
      // under 2^31 bytes)
      const __amdgpu_buffer_rsrc_t rs =
          item_rsrc(pool + (uint32_t)d.x, (uint64_t)((cnt - 1) * ld + (NP - 1) * plane + mt) * 4);
#pragma unroll
      for (int o = 0; o < NP; ++o)
        st.a[o] = buf_load4(rs, (uint32_t)((c * ld + o * plane + s0) * sizeof(float)), 0);
      const int64_t xr = (valid && kl < cnt) ? (int64_t)(ws & kStepRowMask) + c : zrow;
      st.b = __builtin_bit_cast(
          float, __builtin_amdgcn_raw_buffer_load_b32(wrs, (uint32_t)((xr * NB + il) * 4), 0, 0));
    };
    MfsStep<NP> buf[D];
    int4 rec[D];
    // the prologue issues steps 0 .. D-1 in the order the loop re-issues
    // them: the wait counters at the loop header are merged from both
    // paths, and a reordered prologue made the header wait for every load
    // in flight instead of step 0's
#pragma unroll
    for (int j = 0; j < D; ++j) {
      load(record(sb + j), sb + j < se, buf[j]);
      rec[j] = record(sb + D + j);
      __builtin_amdgcn_sched_barrier(0);
    }
    for (int64_t s = sb; s < se; s += D) {
#pragma unroll
      for (int j = 0; j < D; ++j) {
#pragma unroll
        for (int o = 0; o < NP; ++o)
#pragma unroll
          for (int q = 0; q < 4; ++q)
            acc[o][q] =
                __builtin_amdgcn_mfma_f32_16x16x4f32(buf[j].a[o][q], buf[j].b, acc[o][q], 0, 0, 0);
        load(rec[j], s + j + D < se, buf[j]);
        rec[j] = record(s + j + 2 * D);
        // keeps each step's loads right after its products
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  // deterministic cross-wave sum in FP64 (wave 0 stores, the others add in
  // order); rows the clamp duplicated are taken once
  for (int gg = 0; gg < kW; ++gg) {
    if (g == gg) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 4 * kl + r;
        const int si = 4 * i < (int)mt - 4 ? 4 * i : (int)mt - 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = si + q;
          if (row >= 4 * i && row < mt) {
#pragma unroll
            for (int o = 0; o < NP; ++o) {
              const double v = (double)acc[o][q][r];
              double* p = &red[o][il][row];
              *p = (gg == 0) ? v : *p + v;
            }
          }
        }
      }
    }
    __syncthreads();
  }
  double* const ys[3] = {y1, y2, y3};
  for (int e = threadIdx.x; e < NP * NB * 64; e += 64 * kW) {
    const int o = e / (NB * 64), m = e % (NB * 64);
    const int row = ybs > 0 ? m % 64 : m / NB, b = ybs > 0 ? m / 64 : m % NB;
    if (row < mt)
      ys[o][ybs > 0 ? b * ybs + R0 + row : (R0 + row) * NB + b] = red[o][b][row];
  }
}

// FP32 copies of the two interleaved input vectors (m values each): the
// matrix-core products take their weights in FP32.
__global__ void __launch_bounds__(256) k_inputs_f32(int64_t m, const double* __restrict__ v1,
                                                    const double* __restrict__ v2,
                                                    float* __restrict__ x1, float* __restrict__ x2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  x1[i] = (float)v1[i];
  x2[i] = (float)v2[i];
}

// Phase 1 on FP32 matrix cores for 16 right-hand sides, FP32 storage:
// t[l][b] = sum over the unit's columns c of V^T[c][l] x_b[c], one wavefront
// per unit, its member blocks (and rank chunks of 64) one after another.
// A = V^T (ranks x 4 columns), B = the 4 columns' weights (rounded to FP32),
// D = ranks x 16 right-hand sides. Lane l loads ranks s .. s+3 of column
// c + (l >> 4) with one 16-byte load (the four columns of a step are
// contiguous in V^T, so a load instruction reads 4 k consecutive values);
// element q is the A operand of product q, whose rows are ranks 4 i + q. The
// row group start is clamped to k - 4 (to 0 when k < 4: ranks past the
// chunk then read other data and are not stored). The weights are the FP32
// copies x1 = v1, x2 = v2 (interleaved, 16 per cell); each unit's sum (at
// most 512 columns) is formed in FP32 and stored as FP32 rows of t or of the
// partial buffer.
__global__ void __launch_bounds__(256) k_hmv_phase1_mf16(const int64_t* __restrict__ units,
                                                         int64_t nunits,
                                                         const int64_t* __restrict__ members,
                                                         const float* __restrict__ pool,
                                                         const float* __restrict__ x1,
                                                         const float* __restrict__ x2,
                                                         float* __restrict__ t,
                                                         float* __restrict__ part) {
  constexpr int NB = 16;
  const int64_t u = xcd_order(blockIdx.x, gridDim.x) * (blockDim.x / 64) + wave_id();
  if (u >= nunits) return;
  const int lane = threadIdx.x & 63;
  const int kl = lane >> 4, il = lane & 15;
  const int64_t* it = units + u * UN_W;
  const int64_t mem0 = it[UN_MEM], nmem = it[UN_NMEM], cb = it[UN_CB], cc = it[UN_CC];
  const int64_t n = it[UN_N], c0 = it[UN_C0];
  float* dst = it[UN_DIRECT] ? t : part;
  for (int64_t mi = 0; mi < nmem; ++mi) {
    const int64_t* mb = members + (mem0 + mi) * MB_W;
    const int64_t vt = mb[MB_VT], k = mb[MB_K], out = mb[MB_OUT];
    // the member's V^T: 2n x k row-major at pool + vt; the unit reads its
    // columns [cb, cb + cc)
    const __amdgpu_buffer_rsrc_t rs =
        item_rsrc(pool + vt + cb * k, (uint64_t)(cc * k) * sizeof(float));
    const int ldb = (int)(k * (int64_t)sizeof(float));
    for (int64_t l0 = 0; l0 < k; l0 += 64) {
      const int kc = (int)(k - l0 < 64 ? k - l0 : 64);
      const int s0 = (int)l0 + (kc >= 4 ? (4 * il < kc - 4 ? 4 * il : kc - 4) : 0);
      const uint32_t voff = (uint32_t)(s0 * sizeof(float)) + (uint32_t)(kl * ldb);
      f32x4 acc[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
      // four 4-column steps per iteration: all their loads are issued before
      // the products
      int64_t c = 0;
      for (; c + 16 <= cc; c += 16) {
        f32x4 a[4];
        float x[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int64_t g = cb + c + 4 * j + kl;
          x[j] = g < n ? x1[(c0 + g) * NB + il] : x2[(c0 + g - n) * NB + il];
          a[j] = buf_load4(rs, voff, (int)((c + 4 * j) * ldb));
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float bf = x[j];
#pragma unroll
          for (int q = 0; q < 4; ++q)
            acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j][q], bf, acc[q], 0, 0, 0);
        }
      }
      for (; c < cc; c += 4) {
        const int64_t col = c + kl;  // column of the unit (0 .. cc-1)
        const bool in = col < cc;
        const int64_t cl = in ? col : cc - 1;
        const int64_t g = cb + cl;   // stacked column of the block
        const float x = g < n ? x1[(c0 + g) * NB + il] : x2[(c0 + g - n) * NB + il];
        const f32x4 a = buf_load4(rs, voff - (uint32_t)((col - cl) * ldb), (int)(c * ldb));
        const float bf = in ? x : 0.0f;
        const f32x4 av = in ? a : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < 4; ++q)
          acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q], bf, acc[q], 0, 0, 0);
      }
      // lane l, register r holds ranks s_i + q (i = 4 kl + r) of right-hand
      // side il; ranks the clamp duplicated are stored once
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 4 * kl + r;
        const int si = kc >= 4 ? (4 * i < kc - 4 ? 4 * i : kc - 4) : 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int rk = si + q;
          if (rk >= 4 * i && rk < kc) dst[(out + l0 + rk) * NB + il] = acc[q][r];
        }
      }
    }
  }
}

template <typename T, int NB>
static void launch_matvec(const int64_t* units, int64_t nunits, const int64_t* members,
                          const int64_t* red,
                          int64_t nred, const int64_t* tiles, const int64_t* tile_ptr,
                          int64_t ntiles, const int64_t* items, const T* pool, const double* v1,
                          const double* v2, double* t, double* part, double* y1, double* y2,
                          double* y3, int64_t ybs, hipStream_t s) {
  if (nunits > 0)
    k_hmv_phase1<T, NB><<<(unsigned)((nunits + 3) / 4), 256, 0, s>>>(units, nunits, members,
                                                                      pool, v1, v2, t, part);
  if (nred > 0)
    k_hmv_phase1b<NB><<<(unsigned)((nred + 3) / 4), 256, 0, s>>>(red, nred, part, t);
  if (ntiles > 0) {
    if (y3)
      k_hmv_phase2<T, NB, 3><<<(unsigned)ntiles, 64 * kTileWaves * p2_halves<NB>(), 0, s>>>(
          tiles, tile_ptr, items, pool, v1, v2, t, y1, y2, y3, ybs);
    else
      k_hmv_phase2<T, NB, 2><<<(unsigned)ntiles, 64 * kTileWaves * p2_halves<NB>(), 0, s>>>(
          tiles, tile_ptr, items, pool, v1, v2, t, y1, y2, nullptr, ybs);
  }
}

// Scratch of the no-allocation matvec (pq3d_hmat_matvec_noalloc): the
// rank-space vector t and the phase-1 partial rows live in this file-scope
// device array instead of caller buffers. Every GPU that loads the library
// holds its own copy (64 MiB) until the process ends. Host code never uses
// the array's name as an address: the per-device address comes from
// hipGetSymbolAddress, taken once per device with that device current.
constexpr int64_t kScratchDoubles = int64_t(8) << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kScratchMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kScratchMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call per device).
static int scratch_of_current_device(double** out) {
  int dev = -1;
  PQ3D_HIP_CHECK(hipGetDevice(&dev));
  PQ3D_REQUIRE(dev >= 0 && dev < kScratchMaxDevices, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    PQ3D_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PQ3D_REQUIRE(p != nullptr, "no device address for the matvec scratch");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return PQ3D_OK;
}

}  // namespace pq3d

using namespace pq3d;

extern "C" {

// Row widths of every table the H-matrix kernels read, in this order:
// dense block, low-rank block, ACA workspace offsets, matvec item, phase-1
// unit, phase-1b reduction, tile, receiver record, source record, half-space
// source record, phase-1 unit member. Returns the number written.
int pq3d_layout_constants(int64_t* out, int64_t n) {
  const int64_t v[13] = {DB_W, LB_W, WS_W, IT_W, UN_W, RD_W, TL_W, kRecvW, kSrcRec, kHsRec,
                         MB_W, VP_W, PU_W};
  int64_t k = 0;
  for (; k < n && k < 13; ++k) out[k] = v[k];
  return (int)k;
}

int pq3d_hmat_dense_assemble(const int64_t* d_tab, int64_t nblocks, const double* d_recv,
                             const double* d_rec, const double* d_hrec, double mu, double lam,
                             void* d_pool, int dtype_bytes, void* stream) {
  clear_error();
  PQ3D_REQUIRE(nblocks >= 0, "nblocks must be >= 0");
  if (nblocks == 0) return PQ3D_OK;
  PQ3D_REQUIRE(d_tab && d_recv && d_rec && d_pool, "null pointer");
  PQ3D_REQUIRE(dtype_bytes == 4 || dtype_bytes == 8, "dtype_bytes must be 4 or 8");
  PQ3D_REQUIRE(nblocks < (1LL << 24), "too many blocks");
  const double nu = 1.0 / (1.0 + lam / mu) / 2.0;
  hipStream_t s = as_stream(stream);
  const unsigned g = (unsigned)nblocks;
  if (dtype_bytes == 4) {
    if (d_hrec)
      k_dense_assemble<float, true><<<g, 256, 0, s>>>(d_tab, d_recv, d_rec, d_hrec, mu, lam, nu,
                                                      (float*)d_pool);
    else
      k_dense_assemble<float, false><<<g, 256, 0, s>>>(d_tab, d_recv, d_rec, nullptr, mu, lam,
                                                       nu, (float*)d_pool);
  } else {
    if (d_hrec)
      k_dense_assemble<double, true><<<g, 256, 0, s>>>(d_tab, d_recv, d_rec, d_hrec, mu, lam, nu,
                                                       (double*)d_pool);
    else
      k_dense_assemble<double, false><<<g, 256, 0, s>>>(d_tab, d_recv, d_rec, nullptr, mu, lam,
                                                        nu, (double*)d_pool);
  }
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

int pq3d_hmat_aca(const int64_t* d_tab, const int64_t* d_wsoff, int64_t nblocks,
                  const double* d_recv, const double* d_rec, const double* d_hrec, double mu,
                  double lam, double eps,
                  int kmax, double* d_wU, double* d_wV, uint8_t* d_flags, int64_t* d_rank,
                  void* stream) {
  clear_error();
  PQ3D_REQUIRE(nblocks >= 0, "nblocks must be >= 0");
  if (nblocks == 0) return PQ3D_OK;
  PQ3D_REQUIRE(d_tab && d_wsoff && d_recv && d_rec && d_wU && d_wV && d_flags && d_rank,
               "null pointer");
  PQ3D_REQUIRE(kmax >= 1 && kmax <= 4096, "kmax out of range");
  PQ3D_REQUIRE(eps > 0.0, "eps must be > 0");
  PQ3D_REQUIRE(nblocks < (1LL << 24), "too many blocks");
  const double nu = 1.0 / (1.0 + lam / mu) / 2.0;
  if (d_hrec)
    k_aca<true><<<(unsigned)nblocks, kAcaThreads, 0, as_stream(stream)>>>(
        d_tab, d_wsoff, d_recv, d_rec, d_hrec, mu, lam, nu, eps, kmax, d_wU, d_wV, d_flags,
        d_rank);
  else
    k_aca<false><<<(unsigned)nblocks, kAcaThreads, 0, as_stream(stream)>>>(
        d_tab, d_wsoff, d_recv, d_rec, nullptr, mu, lam, nu, eps, kmax, d_wU, d_wV, d_flags,
        d_rank);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

int pq3d_hmat_aca_wide(const int64_t* d_tab, const int64_t* d_wsoff, int64_t nblocks,
                       const double* d_recv, const double* d_rec, const double* d_hrec,
                       double mu, double lam, double eps, int kmax, double* d_wU, double* d_wV,
                       uint8_t* d_flags, int64_t* d_rank, int64_t* d_istate, double* d_dstate,
                       void* stream) {
  clear_error();
  PQ3D_REQUIRE(nblocks >= 0, "nblocks must be >= 0");
  if (nblocks == 0) return PQ3D_OK;
  PQ3D_REQUIRE(d_tab && d_wsoff && d_recv && d_rec && d_wU && d_wV && d_flags && d_rank &&
                   d_istate && d_dstate,
               "null pointer");
  PQ3D_REQUIRE(kmax >= 1 && kmax <= 4096, "kmax out of range");
  PQ3D_REQUIRE(eps > 0.0, "eps must be > 0");
  PQ3D_REQUIRE(nblocks < 65536, "too many blocks for the wide path (grid.y)");
  hipStream_t s = as_stream(stream);
  // block extents on the host: the grids of the row and column steps
  std::vector<int64_t> tab((size_t)nblocks * LB_W);
  PQ3D_HIP_CHECK(hipMemcpyAsync(tab.data(), d_tab, tab.size() * sizeof(int64_t),
                                hipMemcpyDeviceToHost, s));
  PQ3D_HIP_CHECK(hipStreamSynchronize(s));
  int64_t mmax = 0, nmax = 0, kc = 0;
  for (int64_t q = 0; q < nblocks; ++q) {
    const int64_t* b = tab.data() + q * LB_W;
    PQ3D_REQUIRE(b[LB_M] >= 1 && b[LB_N] >= 1, "block %lld is empty", (long long)q);
    mmax = std::max(mmax, b[LB_M]);
    nmax = std::max(nmax, b[LB_N]);
    kc = std::max(kc, std::min<int64_t>(b[LB_K], kmax));
  }
  const double nu = 1.0 / (1.0 + lam / mu) / 2.0;
  unsigned long long* done = reinterpret_cast<unsigned long long*>(d_istate + AW_W * nblocks);
  PQ3D_HIP_CHECK(hipMemsetAsync(done, 0, sizeof(unsigned long long), s));
  const unsigned nb = (unsigned)nblocks;
  k_acaw_init<<<nb, 256, 0, s>>>(d_tab, d_wsoff, kmax, d_flags, d_istate, d_dstate, d_rank,
                                 done);
  PQ3D_LAUNCH_CHECK();
  const dim3 grow(blocks_for(nmax, kAwCols), nb), gcol(blocks_for(mmax, kAwCols), nb);
  const dim3 gcross(blocks_for(kc, 4), nb);
  // a block takes at most kcap rank steps plus 3*64 + 1 zero-row steps
  const int64_t max_steps = kc + 3 * 64 + 2;
  for (int64_t it = 0; it < max_steps; ++it) {
    if (d_hrec)
      k_acaw_row<true><<<grow, kAwCols, 0, s>>>(d_tab, d_wsoff, d_recv, d_rec, d_hrec, mu, lam,
                                                nu, d_wU, d_wV, d_istate);
    else
      k_acaw_row<false><<<grow, kAwCols, 0, s>>>(d_tab, d_wsoff, d_recv, d_rec, nullptr, mu,
                                                 lam, nu, d_wU, d_wV, d_istate);
    k_acaw_pivot<<<nb, kAcaThreads, 0, s>>>(d_tab, d_wsoff, kmax, d_wV, d_flags, d_istate,
                                            d_rank, done);
    if (d_hrec)
      k_acaw_col<true><<<gcol, kAwCols, 0, s>>>(d_tab, d_wsoff, d_recv, d_rec, d_hrec, mu, lam,
                                                nu, d_wU, d_wV, d_istate);
    else
      k_acaw_col<false><<<gcol, kAwCols, 0, s>>>(d_tab, d_wsoff, d_recv, d_rec, nullptr, mu,
                                                 lam, nu, d_wU, d_wV, d_istate);
    k_acaw_cross<<<gcross, 256, 0, s>>>(d_tab, d_wsoff, kmax, d_wU, d_wV, d_istate, d_dstate);
    k_acaw_norm<<<nb, kAcaThreads, 0, s>>>(d_tab, d_wsoff, kmax, eps, d_wU, d_wV, d_flags,
                                           d_istate, d_dstate, d_rank, done);
    PQ3D_LAUNCH_CHECK();
    if (it % 8 == 7 || it + 1 == max_steps) {
      unsigned long long h = 0;
      PQ3D_HIP_CHECK(hipMemcpyAsync(&h, done, sizeof(h), hipMemcpyDeviceToHost, s));
      PQ3D_HIP_CHECK(hipStreamSynchronize(s));
      if (h >= (unsigned long long)nblocks) return PQ3D_OK;
    }
  }
  set_error("wide ACA: blocks still running after %lld steps", (long long)max_steps);
  return PQ3D_EHIP;
}

int pq3d_hmat_aca_pack(const int64_t* d_tab, const int64_t* d_wsoff, int64_t nblocks,
                       const int64_t* d_tiles, const int64_t* d_tile_of_row, const double* d_wU,
                       const double* d_wV, void* d_pool, int dtype_bytes, void* stream) {
  clear_error();
  PQ3D_REQUIRE(nblocks >= 0, "nblocks must be >= 0");
  if (nblocks == 0) return PQ3D_OK;
  PQ3D_REQUIRE(d_tab && d_wsoff && d_tiles && d_tile_of_row && d_wU && d_wV && d_pool,
               "null pointer");
  PQ3D_REQUIRE(dtype_bytes == 4 || dtype_bytes == 8, "dtype_bytes must be 4 or 8");
  PQ3D_REQUIRE(nblocks < (1LL << 24), "too many blocks");
  hipStream_t s = as_stream(stream);
  if (dtype_bytes == 4)
    k_aca_pack<float><<<(unsigned)nblocks, 256, 0, s>>>(d_tab, d_wsoff, d_tiles, d_tile_of_row,
                                                         d_wU, d_wV, (float*)d_pool);
  else
    k_aca_pack<double><<<(unsigned)nblocks, 256, 0, s>>>(d_tab, d_wsoff, d_tiles, d_tile_of_row,
                                                          d_wU, d_wV, (double*)d_pool);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

// y1 = AdotV1, y2 = AdotV2, y3 = dsigma (K part) for nrhs interleaved
// right-hand sides (all FP64, cluster order, element (row, b) of v1, v2 at
// row*nrhs + b). Outputs: the same layout with y_bstride = 0, or element
// (row, b) at b*y_bstride + row (instance-major) with y_bstride > 0.
// d_t: nt*nrhs doubles, d_part: npart*nrhs doubles of scratch. d_y3 may be
// NULL: the dsigma rows (Bs, Bd operators) are then skipped, as the reference
// skips them when the normal traction is fixed (src/QDsim.py:906-909).
int pq3d_hmat_matvec(const int64_t* d_units, int64_t nunits, const int64_t* d_members,
                     const int64_t* d_red, int64_t nred,
                     const int64_t* d_tiles, const int64_t* d_tile_ptr, int64_t ntiles,
                     const int64_t* d_items, const void* d_pool, int dtype_bytes, int nrhs,
                     const double* d_v1, const double* d_v2, double* d_t, double* d_part,
                     double* d_y1, double* d_y2, double* d_y3, int64_t y_bstride,
                     void* stream) {
  clear_error();
  PQ3D_REQUIRE(ntiles >= 0 && nunits >= 0 && nred >= 0, "sizes must be >= 0");
  PQ3D_REQUIRE(y_bstride >= 0, "y_bstride must be >= 0");
  PQ3D_REQUIRE(dtype_bytes == 4 || dtype_bytes == 8, "dtype_bytes must be 4 or 8");
  PQ3D_REQUIRE(d_tiles && d_tile_ptr && d_items && d_pool && d_v1 && d_v2 && d_y1 && d_y2,
               "null pointer");
  PQ3D_REQUIRE(nunits == 0 || (d_units && d_members && d_t), "null pointer");
  PQ3D_REQUIRE(nred == 0 || (d_red && d_part), "null pointer");
  // launch sizes: grid x workgroup size stays below 2^32 threads
  PQ3D_REQUIRE(nunits < (1LL << 26) && nred < (1LL << 26) && ntiles < (1LL << 22),
               "too many work items");
  hipStream_t s = as_stream(stream);
#define PQ3D_MV(NBV)                                                                        \
  do {                                                                                      \
    if (dtype_bytes == 4)                                                                   \
      launch_matvec<float, NBV>(d_units, nunits, d_members, d_red, nred, d_tiles,           \
                                d_tile_ptr, ntiles,                                         \
                                d_items, (const float*)d_pool, d_v1, d_v2, d_t, d_part,     \
                                d_y1, d_y2, d_y3, y_bstride, s);                            \
    else                                                                                    \
      launch_matvec<double, NBV>(d_units, nunits, d_members, d_red, nred, d_tiles,          \
                                 d_tile_ptr, ntiles,                                        \
                                 d_items, (const double*)d_pool, d_v1, d_v2, d_t, d_part,   \
                                 d_y1, d_y2, d_y3, y_bstride, s);                           \
  } while (0)
  switch (nrhs) {
    case 1: PQ3D_MV(1); break;
    case 2: PQ3D_MV(2); break;
    case 4: PQ3D_MV(4); break;
    case 8: PQ3D_MV(8); break;
    case 16: PQ3D_MV(16); break;
    default: pq3d::set_error("nrhs must be 1, 2, 4, 8 or 16"); return PQ3D_EINVAL;
  }
#undef PQ3D_MV
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

// pq3d_hmat_matvec without caller scratch and without any allocation: t
// (nt*nrhs doubles) and the partial rows (npart*nrhs doubles) go into the
// library's static device scratch, at most kScratchDoubles doubles in all;
// larger products are rejected. No element of either needs clearing (phase 1
// or 1b writes every one before phase 2 reads it). Calls that overlap on one
// GPU share the scratch and race: serialise them (one stream, or events).
int pq3d_hmat_matvec_noalloc(const int64_t* d_units, int64_t nunits,
                             const int64_t* d_members, const int64_t* d_red, int64_t nred,
                             const int64_t* d_tiles, const int64_t* d_tile_ptr, int64_t ntiles,
                             const int64_t* d_items, const void* d_pool, int dtype_bytes,
                             int nrhs, const double* d_v1, const double* d_v2, int64_t nt,
                             int64_t npart, double* d_y1, double* d_y2, double* d_y3,
                             int64_t y_bstride, void* stream) {
  clear_error();
  PQ3D_REQUIRE(nt >= 0 && npart >= 0, "nt and npart must be >= 0");
  PQ3D_REQUIRE(nrhs == 1 || nrhs == 2 || nrhs == 4 || nrhs == 8 || nrhs == 16,
               "nrhs must be 1, 2, 4, 8 or 16");
  PQ3D_REQUIRE(nt <= kScratchDoubles && npart <= kScratchDoubles &&
                   (nt + npart) * nrhs <= kScratchDoubles,
               "the product needs (nt + npart) * nrhs = %lld doubles of scratch; the "
               "no-allocation matvec holds at most %lld (use pq3d_hmat_matvec with caller "
               "buffers)",
               (long long)((nt + npart) * nrhs), (long long)kScratchDoubles);
  double* base = nullptr;
  const int rc = scratch_of_current_device(&base);
  if (rc != PQ3D_OK) return rc;
  return pq3d_hmat_matvec(d_units, nunits, d_members, d_red, nred, d_tiles, d_tile_ptr, ntiles,
                          d_items, d_pool, dtype_bytes, nrhs, d_v1, d_v2, base,
                          base + nt * nrhs, d_y1, d_y2, d_y3, y_bstride, stream);
}

// pq3d_hmat_matvec for FP32 storage and 16 right-hand sides with every block
// product on FP32 matrix cores. The inputs are first copied to FP32 (d_x32:
// 2 n 16 floats of scratch, n = rows); phase 1 (V^T x, k_hmv_phase1_mf16)
// sums each unit's columns in FP32 into FP32 rows of t (d_t and d_part are
// used as FP32 arrays; partials of split blocks added in FP64 by phase 1b);
// phase 2 (k_hmv_phase2_mf16) takes those FP32 weights, sums each
// item in FP32 on the matrix cores, accumulates the items of an item group in
// a compensated FP32 pair (compensated != 0) or in plain FP32, and adds the
// four groups in FP64. Same arguments;
// dtype_bytes must be 4 and nrhs 16; d_y3 may be NULL (two row planes);
// every tile must have at least 4 rows.
int pq3d_hmat_matvec_mfma(const int64_t* d_units, int64_t nunits, const int64_t* d_members,
                          const int64_t* d_red, int64_t nred, const int64_t* d_tiles,
                          const int64_t* d_tile_ptr, int64_t ntiles, const int64_t* d_items,
                          const void* d_pool, int dtype_bytes, int nrhs, const double* d_v1,
                          const double* d_v2, double* d_t, double* d_part, double* d_y1,
                          double* d_y2, double* d_y3, int64_t y_bstride, float* d_x32,
                          int64_t n, int compensated, void* stream) {
  clear_error();
  PQ3D_REQUIRE(dtype_bytes == 4, "the matrix-core matvec needs FP32 storage (dtype_bytes 4)");
  PQ3D_REQUIRE(nrhs == 16, "the matrix-core matvec takes 16 right-hand sides");
  PQ3D_REQUIRE(ntiles >= 0 && nunits >= 0 && nred >= 0 && n >= 0, "sizes must be >= 0");
  PQ3D_REQUIRE(y_bstride >= 0, "y_bstride must be >= 0");
  PQ3D_REQUIRE(d_tiles && d_tile_ptr && d_items && d_pool && d_v1 && d_v2 && d_y1 && d_y2 &&
                   d_x32,
               "null pointer");
  PQ3D_REQUIRE(nunits == 0 || (d_units && d_members && d_t), "null pointer");
  PQ3D_REQUIRE(nred == 0 || (d_red && d_part), "null pointer");
  PQ3D_REQUIRE(nunits < (1LL << 26) && nred < (1LL << 26) && ntiles < (1LL << 22) &&
                   n < (1LL << 26),
               "too many work items");
  hipStream_t s = as_stream(stream);
  const float* pool = (const float*)d_pool;
  const int64_t m = n * 16;
  float* x1 = d_x32;
  float* x2 = d_x32 + m;
  float* t32 = (float*)d_t;
  float* part32 = (float*)d_part;
  if (m > 0)
    k_inputs_f32<<<blocks_for(m, 256), 256, 0, s>>>(m, d_v1, d_v2, x1, x2);
  if (nunits > 0)
    k_hmv_phase1_mf16<<<(unsigned)((nunits + 3) / 4), 256, 0, s>>>(
        d_units, nunits, d_members, pool, x1, x2, t32, part32);
  if (nred > 0)
    k_hmv_phase1b<16, float><<<(unsigned)((nred + 3) / 4), 256, 0, s>>>(d_red, nred, part32,
                                                                        t32);
  if (ntiles > 0) {
#define PQ3D_P2MF(NPV, CV, Y3)                                                               \
  k_hmv_phase2_mf16<NPV, CV><<<(unsigned)ntiles, 64 * kMf2Waves, 0, s>>>(                   \
      d_tiles, d_tile_ptr, d_items, pool, x1, x2, t32, d_y1, d_y2, Y3, y_bstride)
    if (d_y3) {
      if (compensated) PQ3D_P2MF(3, true, d_y3);
      else PQ3D_P2MF(3, false, d_y3);
    } else {
      if (compensated) PQ3D_P2MF(2, true, nullptr);
      else PQ3D_P2MF(2, false, nullptr);
    }
#undef PQ3D_P2MF
  }
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

int pq3d_hmat_vpanel_pack(const int64_t* d_tab, int64_t nblocks, const void* d_pool,
                          float* d_vp, void* stream) {
  clear_error();
  PQ3D_REQUIRE(nblocks >= 0, "nblocks must be >= 0");
  if (nblocks == 0) return PQ3D_OK;
  PQ3D_REQUIRE(d_tab && d_pool && d_vp, "null pointer");
  PQ3D_REQUIRE(nblocks < (1LL << 31), "too many blocks");
  k_vpanel_pack<<<(unsigned)nblocks, 256, 0, as_stream(stream)>>>(d_tab, (const float*)d_pool,
                                                                 d_vp);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

int pq3d_hmat_matvec_mfs(const int64_t* d_punits, int64_t npunits, const float* d_vpanels,
                         const int64_t* d_red, int64_t nred, const int32_t* d_steps,
                         const int64_t* d_wptr, const int64_t* d_tiles, int64_t ntiles,
                         const void* d_pool, const double* d_v1, const double* d_v2, float* d_w,
                         int64_t n, int64_t nt, float* d_part, double* d_y1, double* d_y2,
                         double* d_y3, int64_t y_bstride, int p1_depth, int p1_waves, int depth,
                         int waves, void* stream) {
  clear_error();
  PQ3D_REQUIRE(ntiles >= 0 && npunits >= 0 && nred >= 0 && n >= 0 && nt >= 0,
               "sizes must be >= 0");
  PQ3D_REQUIRE(y_bstride >= 0, "y_bstride must be >= 0");
  PQ3D_REQUIRE(depth == 2 || depth == 4 || depth == 6 || depth == 8,
               "depth must be 2, 4, 6 or 8");
  PQ3D_REQUIRE(waves == 1 || waves == 2 || waves == 4, "waves must be 1, 2 or 4");
  PQ3D_REQUIRE(p1_depth == 2 || p1_depth == 4, "p1_depth must be 2 or 4");
  PQ3D_REQUIRE(p1_waves == 2 || p1_waves == 4, "p1_waves must be 2 or 4");
  PQ3D_REQUIRE(d_w && d_y1 && d_y2, "null pointer");
  PQ3D_REQUIRE((d_v1 == nullptr) == (d_v2 == nullptr),
               "d_v1 and d_v2 must both be given or both be NULL");
  PQ3D_REQUIRE(ntiles == 0 || (d_steps && d_wptr && d_tiles && d_pool), "null pointer");
  PQ3D_REQUIRE(npunits == 0 || (d_punits && d_vpanels), "null pointer");
  PQ3D_REQUIRE(nred == 0 || d_part, "null pointer");
  PQ3D_REQUIRE(npunits < (1LL << 31) && nred < (1LL << 26) && ntiles < (1LL << 22) &&
                   n < (1LL << 26),
               "too many work items");
  // weight rows are addressed with 32-bit byte offsets and 28-bit row indices
  PQ3D_REQUIRE((2 * n + nt + 1) * 16 * (int64_t)sizeof(float) < (1LL << 32) &&
                   2 * n + nt < (1LL << 28),
               "weight array too large (2 n + nt = %lld rows)", (long long)(2 * n + nt));
  hipStream_t s = as_stream(stream);
  const float* pool = (const float*)d_pool;
  const int64_t m = n * 16;
  float* x1 = d_w;
  float* x2 = d_w + m;
  float* t32 = d_w + 2 * m;
  const int64_t zrow = 2 * n + nt;
  // NULL inputs: the caller has written the rounded inputs into d_w already
  if (m > 0 && d_v1) k_inputs_f32<<<blocks_for(m, 256), 256, 0, s>>>(m, d_v1, d_v2, x1, x2);
  if (npunits > 0) {
#define PQ3D_MFP(DV, WV)                                                                    \
  if (p1_depth == DV && p1_waves == WV)                                                     \
    k_hmv_phase1_mfp<DV, WV><<<(unsigned)npunits, 64 * WV, 0, s>>>(d_punits, d_vpanels, d_w, \
                                                                   n, zrow, t32, d_part);
    PQ3D_MFP(2, 2) PQ3D_MFP(2, 4) PQ3D_MFP(4, 2) PQ3D_MFP(4, 4)
#undef PQ3D_MFP
  }
  if (nred > 0)
    k_hmv_phase1b<16, float><<<(unsigned)((nred + 3) / 4), 256, 0, s>>>(d_red, nred, d_part, t32);
  if (ntiles > 0) {
    const int4* st = reinterpret_cast<const int4*>(d_steps);
#define PQ3D_MFS(NPV, DV, WV, Y3)                                                          \
  if (depth == DV && waves == WV) {                                                        \
    k_hmv_phase2_mfs<NPV, DV, WV><<<(unsigned)ntiles, 64 * WV, 0, s>>>(                    \
        st, d_wptr, d_tiles, pool, d_w, zrow, d_y1, d_y2, Y3, y_bstride);                  \
  }
#define PQ3D_MFS_ALL(NPV, Y3)                                                              \
  PQ3D_MFS(NPV, 2, 1, Y3) PQ3D_MFS(NPV, 2, 2, Y3) PQ3D_MFS(NPV, 2, 4, Y3)                  \
  PQ3D_MFS(NPV, 4, 1, Y3) PQ3D_MFS(NPV, 4, 2, Y3) PQ3D_MFS(NPV, 4, 4, Y3)                  \
  PQ3D_MFS(NPV, 6, 1, Y3) PQ3D_MFS(NPV, 6, 2, Y3) PQ3D_MFS(NPV, 6, 4, Y3)                  \
  PQ3D_MFS(NPV, 8, 1, Y3) PQ3D_MFS(NPV, 8, 2, Y3) PQ3D_MFS(NPV, 8, 4, Y3)
    if (d_y3) {
      PQ3D_MFS_ALL(3, d_y3)
    } else {
      PQ3D_MFS_ALL(2, nullptr)
    }
#undef PQ3D_MFS_ALL
#undef PQ3D_MFS
  }
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

}  // extern "C"
