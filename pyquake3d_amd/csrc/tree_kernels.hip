// Cluster tree and block partition of the H-matrix on the device, one level
// per launch (the host keeps only the per-level node lists).
//
//  * pq3d_tree_level   every node of one tree level: bounding box of its
//                      points, then, if it has more than `leaf` points, the
//                      split at the box midpoint along the longest extent
//                      (points equal to the midpoint go left) as a stable
//                      partition of its index range; a split that leaves one
//                      side empty makes a leaf (src/Hmatrix.py:133-182)
//  * pq3d_block_level  every (row node, column node) pair of one level of the
//                      block tree: a leaf block if admissible
//                      (min(diam) <= eta * dist, src/Hmatrix.py:58-103) or if
//                      either node is a leaf or has at most min_size points,
//                      else its four child pairs (src/Hmatrix.py:1611-1672).
//                      Each pair carries its path from the root (2 bits per
//                      level), so sorting the leaf blocks by path gives the
//                      depth-first order of the recursion.
// Arithmetic on coordinates is min/max and the same IEEE expressions as the
// host restatement, so trees and partitions are identical to it bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "pq3d_common.h"

namespace pq3d {

constexpr int kTreeThreads = 256;
constexpr int kTreeWaves = kTreeThreads / 64;
constexpr int kMaxBlockDepth = 31;  // 2 path bits per level in 62 bits

__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// One workgroup per node of the level: nodes[2 q] = first index, nodes[2 q + 1]
// = points. box[6 q ..] = (lo, hi) of the node's points; nl[q] = points of the
// left child, or -1 when the node is not split. perm is partitioned in place
// (tmp: scratch of the same size).
__global__ void __launch_bounds__(kTreeThreads) k_tree_level(
    const double* __restrict__ pts, int64_t* __restrict__ perm, int64_t* __restrict__ tmp,
    const int64_t* __restrict__ nodes, int64_t leaf, double* __restrict__ box,
    int64_t* __restrict__ nl_out) {
  __shared__ double red[6][kTreeWaves];
  __shared__ int64_t cnt[kTreeWaves];
  __shared__ double split[2];  // dimension, midpoint
  const int64_t q = blockIdx.x;
  const int64_t s = nodes[2 * q], m = nodes[2 * q + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = tid; i < m; i += kTreeThreads) {
    const int64_t p = perm[s + i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double x = pts[3 * p + k];
      lo[k] = fmin(lo[k], x);
      hi[k] = fmax(hi[k], x);
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lo[k] = wave_min(lo[k]);
    hi[k] = wave_max(hi[k]);
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      red[k][w] = lo[k];
      red[3 + k][w] = hi[k];
    }
  }
  __syncthreads();
  if (tid == 0) {
    double b[6];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      b[k] = red[k][0];
      b[3 + k] = red[3 + k][0];
      for (int v = 1; v < kTreeWaves; ++v) {
        b[k] = fmin(b[k], red[k][v]);
        b[3 + k] = fmax(b[3 + k], red[3 + k][v]);
      }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) box[6 * q + k] = b[k];
    // longest extent, first of equal ones; midpoint as (lo + hi) / 2
    int d = 0;
    double e = b[3] - b[0];
    for (int k = 1; k < 3; ++k)
      if (b[3 + k] - b[k] > e) {
        e = b[3 + k] - b[k];
        d = k;
      }
    split[0] = (double)d;
    split[1] = (b[d] + b[3 + d]) / 2;
  }
  __syncthreads();
  if (m <= leaf) {
    if (tid == 0) nl_out[q] = -1;
    return;
  }
  const int d = (int)split[0];
  const double gamma = split[1];
  // pass 1: points going left
  int64_t c = 0;
  for (int64_t i = tid; i < m; i += kTreeThreads) c += (pts[3 * perm[s + i] + d] <= gamma);
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (lane == 0) cnt[w] = c;
  __syncthreads();
  int64_t nl = 0;
  for (int v = 0; v < kTreeWaves; ++v) nl += cnt[v];
  __syncthreads();
  if (nl == 0 || nl == m) {
    if (tid == 0) nl_out[q] = -1;
    return;
  }
  // pass 2: stable scatter, chunk by chunk; within a chunk a point's rank
  // among the left points is the count of left points on lower lanes
  int64_t cl = 0, cr = 0;  // left / right points placed by earlier chunks
  for (int64_t b0 = 0; b0 < m; b0 += kTreeThreads) {
    const int64_t i = b0 + tid;
    const bool valid = i < m;
    const int64_t p = valid ? perm[s + i] : 0;
    const bool left = valid && pts[3 * p + d] <= gamma;
    const uint64_t bl = __ballot(left);
    const uint64_t bv = __ballot(valid);
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    if (lane == 0) {
      cnt[w] = __popcll(bl);
      red[0][w] = (double)__popcll(bv);
    }
    __syncthreads();
    int64_t wl = 0, wv = 0, tl = 0, tv = 0;
    for (int v = 0; v < kTreeWaves; ++v) {
      if (v < w) {
        wl += cnt[v];
        wv += (int64_t)red[0][v];
      }
      tl += cnt[v];
      tv += (int64_t)red[0][v];
    }
    if (valid) {
      const int64_t rl = wl + __popcll(bl & below);           // left points before i
      const int64_t rv = wv + __popcll(bv & below);           // valid points before i
      const int64_t dst = left ? s + cl + rl : s + nl + cr + (rv - rl);
      tmp[dst] = p;
    }
    cl += tl;
    cr += tv - tl;
    __syncthreads();
  }
  for (int64_t i = tid; i < m; i += kTreeThreads) perm[s + i] = tmp[s + i];
  if (tid == 0) nl_out[q] = nl;
}

// One thread per pair. node[4 a ..] = (first, points, left, right); box[6 a ..]
// = (lo, hi). pairs[3 p ..] = (row node, col node, path). Leaf blocks go to
// leafs[4 r ..] = (row node, col node, admissible, path), child pairs to
// next[3 r ..]; counts[0], counts[1] are their running totals (zero on entry).
__global__ void __launch_bounds__(256) k_block_level(
    const int64_t* __restrict__ pairs, int64_t npairs, const int64_t* __restrict__ node,
    const double* __restrict__ box, double eta, int64_t min_size, int depth,
    int64_t* __restrict__ leafs, int64_t* __restrict__ next,
    unsigned long long* __restrict__ counts) {
  const int64_t pi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pi >= npairs) return;
  const int64_t a = pairs[3 * pi], b = pairs[3 * pi + 1], path = pairs[3 * pi + 2];
  const double* la = box + 6 * a;
  const double* lb = box + 6 * b;
  double ea[3], eb[3], g[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    ea[k] = la[3 + k] - la[k];
    eb[k] = lb[3 + k] - lb[k];
    const double g1 = fmax(0.0, lb[k] - la[3 + k]);
    const double g2 = fmax(0.0, la[k] - lb[3 + k]);
    g[k] = g1 * g1 + g2 * g2;
  }
  const double da = sqrt(ea[0] * ea[0] + ea[1] * ea[1] + ea[2] * ea[2]);
  const double db = sqrt(eb[0] * eb[0] + eb[1] * eb[1] + eb[2] * eb[2]);
  const double gap = g[0] + g[1] + g[2];
  const bool adm = fmin(da, db) <= eta * sqrt(gap);
  const int64_t* na = node + 4 * a;
  const int64_t* nb = node + 4 * b;
  if (adm || na[1] <= min_size || nb[1] <= min_size || na[2] < 0 || nb[2] < 0) {
    const unsigned long long r = atomicAdd(&counts[0], 1ull);
    leafs[4 * r] = a;
    leafs[4 * r + 1] = b;
    leafs[4 * r + 2] = adm ? 1 : 0;
    leafs[4 * r + 3] = path;
    return;
  }
  // children in the recursion's visiting order: (la, lb), (la, rb), (ra, lb), (ra, rb)
  const unsigned long long r = atomicAdd(&counts[1], 4ull);
  const int shift = 2 * (kMaxBlockDepth - 1 - depth);
  const int64_t ca[4] = {na[2], na[2], na[3], na[3]};
  const int64_t cb[4] = {nb[2], nb[3], nb[2], nb[3]};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    next[3 * (r + c)] = ca[c];
    next[3 * (r + c) + 1] = cb[c];
    next[3 * (r + c) + 2] = path | ((int64_t)c << shift);
  }
}

}  // namespace pq3d

using namespace pq3d;

extern "C" {

int pq3d_tree_level(const double* d_pts, int64_t n, int64_t* d_perm, int64_t* d_tmp,
                    const int64_t* d_nodes, int64_t nnodes, int64_t leaf, double* d_box,
                    int64_t* d_nl, void* stream) {
  clear_error();
  PQ3D_REQUIRE(n >= 0 && nnodes >= 0, "sizes must be >= 0");
  if (nnodes == 0) return PQ3D_OK;
  PQ3D_REQUIRE(d_pts && d_perm && d_tmp && d_nodes && d_box && d_nl, "null pointer");
  PQ3D_REQUIRE(leaf >= 1, "leaf must be >= 1");
  PQ3D_REQUIRE(nnodes < ((int64_t)1 << 24), "too many nodes in one level (at most 2^24 - 1)");
  k_tree_level<<<(unsigned)nnodes, kTreeThreads, 0, as_stream(stream)>>>(
      d_pts, d_perm, d_tmp, d_nodes, leaf, d_box, d_nl);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

int pq3d_block_level(const int64_t* d_pairs, int64_t npairs, const int64_t* d_node,
                     const double* d_box, double eta, int64_t min_size, int depth,
                     int64_t* d_leafs, int64_t* d_next, unsigned long long* d_counts,
                     void* stream) {
  clear_error();
  PQ3D_REQUIRE(npairs >= 0, "npairs must be >= 0");
  if (npairs == 0) return PQ3D_OK;
  PQ3D_REQUIRE(d_pairs && d_node && d_box && d_leafs && d_next && d_counts, "null pointer");
  PQ3D_REQUIRE(depth >= 0 && depth < kMaxBlockDepth - 1, "block tree deeper than %d levels",
               kMaxBlockDepth - 1);
  PQ3D_REQUIRE(npairs <= ((int64_t)1 << 31), "too many pairs in one level (at most 2^31)");
  k_block_level<<<blocks_for(npairs, 256), 256, 0, as_stream(stream)>>>(
      d_pairs, npairs, d_node, d_box, eta, min_size, depth, d_leafs, d_next, d_counts);
  PQ3D_LAUNCH_CHECK();
  return PQ3D_OK;
}

}  // extern "C"
