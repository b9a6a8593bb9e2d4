/*
 * pyquake3d_amd -- C-ABI of the MI355X (gfx950) hot path of PyQuake3D.
 *
 * Library: pyquake3d_amd/lib/libpq3d.so (built by `python -m pyquake3d_amd.build`).
 *
 * Conventions
 *   - Functions named pq3d_* return 0 on success or a negative status
 *     (PQ3D_EINVAL -1, PQ3D_EHIP -2, PQ3D_ENOTSUP -3, PQ3D_ERANGE -4) and leave
 *     a message in pq3d_last_error() (thread-local).
 *   - Pointers named d_* are device pointers (hipMalloc / torch CUDA tensors);
 *     `stream` is a hipStream_t (NULL = default stream). These calls are
 *     asynchronous on `stream`.
 *   - The two reference symbols take host pointers, are synchronous, and keep
 *     the reference's void signature.
 *   - Stress/strain order: xx, yy, zz, xy, xz, yz. Coordinates in metres,
 *     moduli in Pa, slip in metres. Triangles are 9 doubles: P1, P2, P3.
 *   - Traction operator order: A1s, A2s, Bs, A1d, A2d, Bd (strike slip Ss=1
 *     then dip slip Ds=1; traction along ev1, ev2 and ev3 of the receiver).
 */
#ifndef PYQUAKE3D_AMD_H
#define PYQUAKE3D_AMD_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ library */

/* Message of the last failing pq3d_* call on this thread ("" if none). */
const char* pq3d_last_error(void);
/* ABI version of this library. */
int pq3d_version(void);

/* ------------------------------------------- reference drop-in symbols */

/* Replaces `TDstressFS_C` (reference src/TDstressFS_C.cpp:2048-2095, bound by
 * ctypes at src/Hmatrix.py:25-38): stress and strain at n receivers
 * (X[i],Y[i],Z[i]) from one triangle (P1,P2,P3, 3 doubles each) with Burgers
 * vector (Ss, Ds, Ts), in the full space or, with jud_halfspace, the half
 * space z <= 0. Outputs n x 6 row-major. On an error the outputs become NaN
 * and pq3d_last_error() says why. */
void TDstressFS_C(const double* X, const double* Y, const double* Z, size_t n,
                  const double* P1, const double* P2, const double* P3, double Ss, double Ds,
                  double Ts, double mu, double lambda_, bool jud_halfspace, double* stress,
                  double* strain);

/* Replaces `TDstressEachSourceAtReceiver_C` (reference
 * src/TDstressFS_C.cpp:2098-2143, bound at src/Hmatrix.py:40-49): one receiver
 * (x,y,z) and n_sources triangles given as flat n_sources x 3 arrays. Outputs
 * n_sources x 6 row-major. Half space and errors as above. */
void TDstressEachSourceAtReceiver_C(double x, double y, double z, const double* P1_flat,
                                    const double* P2_flat, const double* P3_flat,
                                    size_t n_sources, double Ss, double Ds, double Ts, double mu,
                                    double lambda_, bool jud_halfspace, double* stress_out,
                                    double* strain_out);

/* ---------------------------------------------- device TDE kernels */

/* Per-triangle setup shared by all receivers, computed on the HOST with the C
 * library's acos/cos/sin (so that it is bitwise what the reference computes
 * per call, src/TDstressFS_C.cpp:292-395): tri (n x 9, host) -> rec (n x 32
 * doubles, host). Upload rec to the device for the calls below. */
int pq3d_tde_source_records(const double* tri, int64_t n, double* rec);

/* Device form of TDstressFS_C: n receivers x one source record d_rec1. */
int pq3d_tde_stress_points(const double* d_X, const double* d_Y, const double* d_Z, int64_t n,
                           const double* d_rec1, double Ss, double Ds, double Ts, double mu,
                           double lambda_, double* d_stress, double* d_strain, void* stream);

/* Device form of TDstressEachSourceAtReceiver_C: one receiver x ns records. */
int pq3d_tde_stress_sources(double x, double y, double z, const double* d_rec, int64_t ns,
                            double Ss, double Ds, double Ts, double mu, double lambda_,
                            double* d_stress, double* d_strain, void* stream);

/* Traction operators of an m x n block (replaces the column evaluators
 * calc_stressfunc_fromCS/CD, src/Hmatrix.py:425-511, the row evaluators
 * calc_stressfunc_fromC_rowS/rowD, :318-422, and GetTtstress, :295-313).
 * d_recv: receiver records (N x 12: centroid, ev1, ev2, ev3);
 * d_rec: source records (N x 32). Receiver i is d_rows[i] (or row0+i when
 * d_rows is NULL), source j is d_cols[j] (or col0+j).
 * Output element (op, i, j) goes to d_out[op*plane + i*si + j*sj]; out_len is
 * the number of elements of d_out (checked). FP64 and FP32-output variants. */
int pq3d_tde_traction_block(const double* d_recv, const double* d_rec, const int64_t* d_rows,
                            int64_t row0, int64_t m, const int64_t* d_cols, int64_t col0,
                            int64_t n, double mu, double lambda_, double* d_out, int64_t si,
                            int64_t sj, int64_t plane, int64_t out_len, void* stream);
int pq3d_tde_traction_block_f32(const double* d_recv, const double* d_rec,
                                const int64_t* d_rows, int64_t row0, int64_t m,
                                const int64_t* d_cols, int64_t col0, int64_t n, double mu,
                                double lambda_, float* d_out, int64_t si, int64_t sj,
                                int64_t plane, int64_t out_len, void* stream);

/* ------------------------------------------------ half space (z <= 0)
 * TDstressHS (src/TDstressFS_C.cpp:1911-1999): main + free-surface
 * correction (TDstress_HarFunc, :1766-1908, with the reference's reuse of
 * the previous side's values for a vertical side) + image. The correction
 * strain is the gradient of the correction displacement
 * (src/DH_greenfunction.py:313-399) by forward-mode differentiation: it
 * agrees with the reference to rounding, not bit for bit.
 * pq3d_tde_hs_records: host, 104 doubles per triangle (tri n x 9). */
int pq3d_tde_hs_records(const double* tri, int64_t n, double* hrec);
/* TDstressFS_C with jud_halfspace = true, device pointers. */
int pq3d_tde_stress_points_hs(const double* d_X, const double* d_Y, const double* d_Z,
                              int64_t n, const double* d_rec1, const double* d_hrec1, double Ss,
                              double Ds, double Ts, double mu, double lambda_, double* d_stress,
                              double* d_strain, void* stream);
/* TDstressEachSourceAtReceiver_C with jud_halfspace = true, device pointers. */
int pq3d_tde_stress_sources_hs(double x, double y, double z, const double* d_rec,
                               const double* d_hrec, int64_t ns, double Ss, double Ds, double Ts,
                               double mu, double lambda_, double* d_stress, double* d_strain,
                               void* stream);
/* pq3d_tde_traction_block in the half space (the evaluators of
 * src/Hmatrix.py:318-511 with halfspace_jud = True). */
int pq3d_tde_traction_block_hs(const double* d_recv, const double* d_rec, const double* d_hrec,
                               const int64_t* d_rows, int64_t row0, int64_t m,
                               const int64_t* d_cols, int64_t col0, int64_t n, double mu,
                               double lambda_, double* d_out, int64_t si, int64_t sj,
                               int64_t plane, int64_t out_len, void* stream);
int pq3d_tde_traction_block_hs_f32(const double* d_recv, const double* d_rec,
                                   const double* d_hrec, const int64_t* d_rows, int64_t row0,
                                   int64_t m, const int64_t* d_cols, int64_t col0, int64_t n,
                                   double mu, double lambda_, float* d_out, int64_t si,
                                   int64_t sj, int64_t plane, int64_t out_len, void* stream);

/* ------------------------------------------------------------ displacement
 * Displacement of a triangular dislocation, full space (TDdispFS,
 * src/DH_greenfunction.py:99-220) and half space (TDdispHS, :402-443: main
 * + free-surface correction + image). The reference binds no C symbol for
 * these (they are numpy functions); nu is Poisson's ratio. A side whose
 * direction is vertical adds no correction, as in the published method. */
/* TDdispFS / TDdispHS: n points x one triangle (d_hrec1 null -> full space).
 * d_disp n x 3 row-major (east, north, up); NaN at points on a side of the
 * triangle in its own plane, as the reference. */
int pq3d_tde_disp_points(const double* d_X, const double* d_Y, const double* d_Z, int64_t n,
                         const double* d_rec1, const double* d_hrec1, double Ss, double Ds,
                         double Ts, double nu, double* d_disp, void* stream);
/* get_sumS's displacement (src/QDsim.py:25-53) with per-triangle slip:
 * d_out[3 i + c] = sum_j u_c(d_obs[3 i .. 3 i + 2]; triangle j, d_ss[j],
 * d_ds[j], d_ts ? d_ts[j] : 0). d_obs nobs x 3 (nobs <= 2^29); d_hrec null
 * -> full space. Deterministic: the same inputs give the same bits. */
int pq3d_tde_disp_sum(const double* d_obs, int64_t nobs, const double* d_rec,
                      const double* d_hrec, int64_t ns, const double* d_ss, const double* d_ds,
                      const double* d_ts, double nu, double* d_out, void* stream);

/* ------------------------------------------------------- cluster tree
 * Cluster tree and block partition on the device, one level per call
 * (src/Hmatrix.py:133-182 build_block_tree, :58-103 admissibility,
 * :1611-1672 create_recursive_blocks). The same comparisons and IEEE
 * expressions as the host restatement: identical trees and partitions. */
/* One tree level. d_pts n x 3 (original order); d_perm (n) holds the current
 * order and is partitioned in place (d_tmp: n scratch). d_nodes nnodes x 2 =
 * (first index, points). Out: d_box nnodes x 6 = (lo, hi) of each node's
 * points; d_nl[q] = points of the left child, or -1 when node q is not split
 * (at most `leaf` points, or one side empty). */
int pq3d_tree_level(const double* d_pts, int64_t n, int64_t* d_perm, int64_t* d_tmp,
                    const int64_t* d_nodes, int64_t nnodes, int64_t leaf, double* d_box,
                    int64_t* d_nl, void* stream);
/* One block-tree level. d_pairs npairs x 3 = (row node, col node, path);
 * d_node nodes x 4 = (first, points, left, right; -1 for a leaf node);
 * d_box nodes x 6. Admissible pairs and pairs with a leaf node or a node of
 * at most min_size points go to d_leafs (x 4: row, col, admissible, path),
 * the others put their four children in d_next (x 3); d_counts[0..1] (zero
 * on entry) count both. Capacity: npairs leafs, 4 npairs next. The path of a
 * child has its index 0..3 in bits 2 (30 - depth) (order (ll, lr, rl, rr));
 * depth < 30. */
int pq3d_block_level(const int64_t* d_pairs, int64_t npairs, const int64_t* d_node,
                     const double* d_box, double eta, int64_t min_size, int depth,
                     int64_t* d_leafs, int64_t* d_next, unsigned long long* d_counts,
                     void* stream);

/* ------------------------------------------------------------ H-matrix
 * Index space: cluster order (the cluster tree's permutation); blocks are
 * contiguous ranges. Pool element type: float (dtype_bytes 4) or double (8).
 * Stacked block operator: K[c*m+i, s*n+j] = op(c,s)[i,j] with
 * op(0,0)=A1s op(1,0)=A2s op(2,0)=Bs op(0,1)=A1d op(1,1)=A2d op(2,1)=Bd. */

/* Row widths (int64 elements) of the tables the kernels below read, in the
 * order: dense block, low-rank block, ACA workspace offsets, matvec item,
 * phase-1 unit, phase-1b reduction, tile, receiver record, source record,
 * half-space source record, phase-1 unit member, V panel packing row, panel
 * phase-1 unit. Host-only; returns the number of values written (at most
 * n). */
int pq3d_layout_constants(int64_t* out, int64_t n);

/* Near-field blocks (replaces ACA_worker's dense branch with
 * calc_stressfunc_fromCS/CD, src/Hmatrix.py:1049-1064, :425-511).
 * d_tab: nblocks x 6 int64 (r0, m, c0, n, pool base, unused); block data is
 * six column-major m x n planes at d_pool[base + op*m*n + j*m + i].
 * d_hrec: half-space records (pq3d_tde_hs_records) for the half space, or
 * NULL for the full space (also for pq3d_hmat_aca). */
int pq3d_hmat_dense_assemble(const int64_t* d_tab, int64_t nblocks, const double* d_recv,
                             const double* d_rec, const double* d_hrec, double mu,
                             double lambda_, void* d_pool, int dtype_bytes, void* stream);

/* Far-field ACA (replaces ACA_plus_full_traceS/D, src/Hmatrix.py:581-1011,
 * with a relative Frobenius stop): one workgroup per block of the stacked
 * 3m x 2n operator; entries are evaluated in the kernel. d_tab: nblocks x 8
 * int64 (r0, m, c0, n, -, -, rank cap, -); d_wsoff: nblocks x 4 int64
 * offsets into d_wU (cap*3m doubles per block), d_wV (cap*2n) and d_flags
 * (3m+2n bytes). d_rank[b] = rank, or -1 if the cap was reached. */
int pq3d_hmat_aca(const int64_t* d_tab, const int64_t* d_wsoff, int64_t nblocks,
                  const double* d_recv, const double* d_rec, const double* d_hrec, double mu,
                  double lambda_, double eps, int kmax, double* d_wU, double* d_wV,
                  uint8_t* d_flags, int64_t* d_rank, void* stream);

/* The same ACA for blocks too large for one workgroup (same arguments and
 * results as pq3d_hmat_aca; the factors and ranks are bitwise equal): every
 * rank step is a few launches over all the blocks, with the row and column
 * residuals spread over workgroups of 256 columns / receivers. Returns after
 * every block has finished (the host checks every 8 steps). Scratch from the
 * caller: d_istate nblocks*8 + 1 int64, d_dstate nblocks*(kmax + 1) doubles.
 * nblocks < 65536. */
int pq3d_hmat_aca_wide(const int64_t* d_tab, const int64_t* d_wsoff, int64_t nblocks,
                       const double* d_recv, const double* d_rec, const double* d_hrec,
                       double mu, double lambda_, double eps, int kmax, double* d_wU,
                       double* d_wV, uint8_t* d_flags, int64_t* d_rank, int64_t* d_istate,
                       double* d_dstate, void* stream);

/* Copy converged factors into the pool (the storage half of ACA_worker,
 * src/Hmatrix.py:1014-1066): U (3m x k) at d_tab[b][4] in one panel per leaf
 * cluster of the block's rows -- the leaf with row offset ro and ls rows in
 * the block holds element (o*m + i, l) at 3*k*ro + o*k*ls + l*ls + (i - ro);
 * V transposed (2n x k, row-major) at d_tab[b][5]; k = d_tab[b][6].
 * d_tiles: leaf clusters as (first row, rows) int64 pairs; d_tile_of_row:
 * leaf index of every cluster-order row. */
int pq3d_hmat_aca_pack(const int64_t* d_tab, const int64_t* d_wsoff, int64_t nblocks,
                       const int64_t* d_tiles, const int64_t* d_tile_of_row, const double* d_wU,
                       const double* d_wV, void* d_pool, int dtype_bytes, void* stream);

/* Six-operator H-matvec for nrhs right-hand sides (replaces the six
 * blocks_process_MVM calls of one step, src/Hmatrix.py:1445-1538 via
 * src/QDsim.py:906-917): y1 = A1s v1 + A1d v2, y2 = A2s v1 + A2d v2,
 * y3 = Bs v1 + Bd v2, FP64 inputs interleaved by instance (element (row, b)
 * at row*nrhs + b), FP64 accumulation; nrhs in {1, 2, 4, 8, 16}. Outputs in
 * the same layout (y_bstride = 0) or instance-major, element (row, b) at
 * b*y_bstride + row (y_bstride > 0, normally the row count). Low-rank V
 * factors are stored transposed (2n x k), U in leaf panels. Work lists
 * (phase-1 units and their member blocks, phase-1b reductions, tiles, items)
 * as built by
 * pyquake3d_amd.hmatrix.build_worklists; d_t: nt*nrhs and d_part:
 * npart*nrhs doubles of scratch. d_y3 may be NULL: the Bs and Bd operators
 * are then neither read nor applied, as in the reference's step with a fixed
 * normal traction (src/QDsim.py:906-909, Fix_Tn). */
int pq3d_hmat_matvec(const int64_t* d_units, int64_t nunits, const int64_t* d_members,
                     const int64_t* d_red, int64_t nred,
                     const int64_t* d_tiles, const int64_t* d_tile_ptr, int64_t ntiles,
                     const int64_t* d_items, const void* d_pool, int dtype_bytes, int nrhs,
                     const double* d_v1, const double* d_v2, double* d_t, double* d_part,
                     double* d_y1, double* d_y2, double* d_y3, int64_t y_bstride,
                     void* stream);

/* pq3d_hmat_matvec without scratch arguments and without any allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The rank-space vector t and the partial
 * rows (nt*nrhs and npart*nrhs doubles, sizes from the work lists) live in
 * the library's static device scratch of 8 Mi doubles (64 MiB per GPU, held
 * until the process ends); a product needing more returns PQ3D_EINVAL.
 * Calls that overlap on one GPU share that scratch and race. */
int pq3d_hmat_matvec_noalloc(const int64_t* d_units, int64_t nunits,
                             const int64_t* d_members, const int64_t* d_red, int64_t nred,
                             const int64_t* d_tiles, const int64_t* d_tile_ptr, int64_t ntiles,
                             const int64_t* d_items, const void* d_pool, int dtype_bytes,
                             int nrhs, const double* d_v1, const double* d_v2, int64_t nt,
                             int64_t npart, double* d_y1, double* d_y2, double* d_y3,
                             int64_t y_bstride, void* stream);

/* pq3d_hmat_matvec with every block product on FP32 matrix cores (same
 * replaced reference path, src/Hmatrix.py:1445-1538): FP32 storage
 * (dtype_bytes 4), nrhs 16 and tiles of at least 4 rows only. The inputs are
 * rounded to FP32 into d_x32 (2 n 16 floats of scratch, n = rows); d_t and
 * d_part serve as FP32 arrays (nt*16 and npart*16 floats). Phase 1 sums each
 * phase-1 unit's columns (at most 512) in FP32, split blocks' partials in
 * FP64; phase 2 sums each item in FP32 on the matrix cores and adds the
 * items of an item group in compensated FP32 (hi + lo pairs, TwoSum) when
 * compensated != 0, else in plain FP32; the four item groups in FP64;
 * outputs FP64. d_y3 may be NULL (two planes). */
int pq3d_hmat_matvec_mfma(const int64_t* d_units, int64_t nunits, const int64_t* d_members,
                          const int64_t* d_red, int64_t nred, const int64_t* d_tiles,
                          const int64_t* d_tile_ptr, int64_t ntiles, const int64_t* d_items,
                          const void* d_pool, int dtype_bytes, int nrhs, const double* d_v1,
                          const double* d_v2, double* d_t, double* d_part, double* d_y1,
                          double* d_y2, double* d_y3, int64_t y_bstride, float* d_x32,
                          int64_t n, int compensated, void* stream);

/* V panels of the matrix-core product (pq3d_hmat_matvec_mfs), packed from an
 * FP32 pool. d_tab: nblocks x 6 int64 per low-rank block (its V^T base in
 * the pool (2n x k, row-major), k, n, panel base in d_vp, panel rank K,
 * first rank of the block in the panel). A panel holds every low-rank block
 * of one source cluster side by side (2n x K), chunk-major: ranks
 * [64 j, 64 j + kc) form a contiguous 2n x kc row-major chunk at
 * base + 2n 64 j. No counterpart in the reference (its blocks keep separate
 * factors, src/Hmatrix.py:1014-1066); it is the storage the product below
 * streams. */
int pq3d_hmat_vpanel_pack(const int64_t* d_tab, int64_t nblocks, const void* d_pool,
                          float* d_vp, void* stream);

/* The 16-right-hand-side product on FP32 matrix cores (replaces the same
 * reference path, src/Hmatrix.py:1445-1538: the six blocks_process_MVM calls
 * of a step for 16 simulations sharing the operator). FP32 storage, tiles
 * of at least 4 rows.
 * Weights d_w: (2 n + nt + 1) rows of 16 floats: the inputs rounded to FP32
 * (written by the call from d_v1, d_v2, or, with both NULL, already written
 * by the caller, e.g. with pq3d_deficit_w32), t (written by phase 1) and a
 * last row of zeros (the caller keeps it zero).
 * Phase 1 (t = V^T x): d_punits, npunits x 8 int64 units over the V panels
 * d_vpanels (pq3d_hmat_vpanel_pack): (chunk element of the unit's first
 * column, ranks kc <= 64, output row, first stacked column, columns, the
 * panel's cluster n and c0, direct flag); a unit is p1_waves (2 or 4)
 * wavefronts with p1_depth (2 or 4) 4-column steps in flight each; direct
 * units write t rows, the others partial rows of d_part (npart*16 floats)
 * summed by the d_red rows as in pq3d_hmat_matvec (t rows numbered panel by
 * panel).
 * Phase 2: d_steps, one 4 x int32 record per 4-column step of every item
 * (pool element of the step's first column at its tile's first row, column
 * stride, plane stride, (columns - 1) << 28 | weight row), grouped by tile;
 * d_wptr: ntiles*waves + 1 offsets, the contiguous step range of wavefront g
 * of tile q at [wptr[q*waves + g], wptr[q*waves + g + 1]); depth (2, 4, 6 or 8)
 * steps in flight per wavefront, waves (1, 2 or 4) wavefronts per tile.
 * FP32 sums per wavefront, FP64 across wavefronts; outputs FP64 as in
 * pq3d_hmat_matvec. d_y3 may be NULL (two planes). */
int pq3d_hmat_matvec_mfs(const int64_t* d_punits, int64_t npunits, const float* d_vpanels,
                         const int64_t* d_red, int64_t nred, const int32_t* d_steps,
                         const int64_t* d_wptr, const int64_t* d_tiles, int64_t ntiles,
                         const void* d_pool, const double* d_v1, const double* d_v2, float* d_w,
                         int64_t n, int64_t nt, float* d_part, double* d_y1, double* d_y2,
                         double* d_y3, int64_t y_bstride, int p1_depth, int p1_waves, int depth,
                         int waves, void* stream);

/* ---------------------------------------------------------- time step
 * Cells may belong to n_inst independent instances (d_inst[i], or NULL for
 * one instance); d_h, d_cellpar (V0, mu, Cs per instance), d_err (4 per
 * instance) and d_vmax are indexed by instance. Error and max slots are
 * uint64 bit patterns of non-negative doubles (zeroed by the calls). */

/* Slip-rate deficit (src/QDsim.py:902-903): x = V - Vpl (cos, sin)(rake0). */
int pq3d_deficit(int64_t n, const double* d_slipv1, const double* d_slipv2,
                 const double* d_pl1, const double* d_pl2, double* d_x1, double* d_x2,
                 void* stream);

/* The same for nb instances of n cells stored instance-major (cell c of
 * instance b at b*n + c), written interleaved for the multi-RHS H-matvec:
 * x[c*nb + b] = V[b*n + c] - Vpl[b*n + c]. */
int pq3d_deficit_interleave(int64_t n, int64_t nb, const double* d_slipv1,
                            const double* d_slipv2, const double* d_pl1, const double* d_pl2,
                            double* d_x1, double* d_x2, void* stream);

/* The same interleaved deficit written straight into the FP32 weight rows of
 * pq3d_hmat_matvec_mfs: d_w[c*nb + b] = x1 and d_w[n*nb + c*nb + b] = x2 of
 * cell c, instance b, each the FP64 difference rounded to FP32 (what the
 * product's own input conversion would give). */
int pq3d_deficit_w32(int64_t n, int64_t nb, const double* d_slipv1, const double* d_slipv2,
                     const double* d_pl1, const double* d_pl2, float* d_w, void* stream);

/* dsigma/dt = K part + normal loading (src/QDsim.py:911-913); fix_tn: loading only. */
int pq3d_add_loading(int64_t n, double* d_dsig, const double* d_load, int fix_tn,
                     void* stream);

/* Traction-rate RHS (replaces QDsim.derivative, src/QDsim.py:760-813):
 * d_out = 4 x n (dpsi/dt, dsigma/dt, dtau1/dt, dtau2/dt in MPa/s). */
int pq3d_qd_rhs(int64_t n, const double* d_Tno, const double* d_Tt1o, const double* d_Tt2o,
                const double* d_state, const double* d_a, const double* d_b, const double* d_dc,
                const double* d_f0, const double* d_AdotV1, const double* d_AdotV2,
                const double* d_dsigmadt, const double* d_shear_loading, const double* d_slipv,
                double V0, double mu, double Cs, double* d_out, void* stream);

/* One fused Dormand-Prince 5(4) attempt of step d_h[inst] (replaces
 * RungeKutte_solve_Dormand_Prince, src/QDsim.py:1001-1134): candidate state
 * to d_y*, error maxima (psi, tau1, tau2, sigma) to d_err. a63 is the
 * tableau's a63 (the reference uses -46732/5247, the published value is
 * +46732/5247). An instance with d_h[inst] < 0 is left out: its d_y* cells
 * keep their values (the retry rounds of an ensemble skip the instances
 * that have already accepted). */
int pq3d_rk45_attempt(int64_t n, const int32_t* d_inst, const double* d_h, const double* d_Tno,
                      const double* d_Tt1o, const double* d_Tt2o, const double* d_state,
                      const double* d_a, const double* d_b, const double* d_dc,
                      const double* d_f0, const double* d_AdotV1, const double* d_AdotV2,
                      const double* d_dsigmadt, const double* d_shear_loading,
                      const double* d_slipv, const double* d_cellpar, double* d_yTno,
                      double* d_yTt1o, double* d_yTt2o, double* d_ystate,
                      unsigned long long* d_err, double a63, int64_t n_inst, void* stream);

/* Accept the candidate (src/QDsim.py:960-993) for instances with
 * d_accept[inst] == 1 (all when NULL): state, slip rates from the
 * regularized law per component, rake, V floor 1e-30, slip; d_vmax = max V. */
int pq3d_rk45_commit(int64_t n, const int32_t* d_inst, const double* d_h,
                     const uint8_t* d_accept, const double* d_a, const double* d_cellpar,
                     const double* d_yTno, const double* d_yTt1o, const double* d_yTt2o,
                     const double* d_ystate, double* d_Tno, double* d_Tt1o, double* d_Tt2o,
                     double* d_state, double* d_slipv1, double* d_slipv2, double* d_slipv,
                     double* d_rake, double* d_slip1, double* d_slip2,
                     unsigned long long* d_vmax, int64_t n_inst, void* stream);

/* Device-driven form of the accept/reject loop of simu_forward
 * (src/QDsim.py:924-958): `rounds` times, one attempt of every instance with
 * d_flag 0 at d_h, then the reference's decision on device (accept: flag 1,
 * d_dtnext = min(hn1, hn2, 1.5 h); reject: h = max(0.5 h, min(hn1, hn2));
 * flag 2 and d_status 1 when h < 1e-15 or after 20 rejections). No host
 * synchronisation; the operator products must be current for the step.
 * d_lstate: NULL, or one learned-state record per instance (PQ3D_LS_W
 * doubles: learned flag, last accepted max(c1, c2), last dtnext/h, dtnext
 * base, six features). With it, an accepted step stores its max(c1, c2) and
 * dtnext/h, and a rejected attempt of a step that began with a learned
 * proposal takes the PI fallback h 0.9 (1/err)^0.14 prev^0.08 within
 * [0.2 h, 0.9 h] (no reference counterpart; the reference rule otherwise). */
int pq3d_rk45_rounds(int64_t n, const int32_t* d_inst, double* d_h, uint8_t* d_flag,
                     int32_t* d_nrj, double* d_dtnext, int32_t* d_status, int rounds,
                     double tol, const double* d_Tno, const double* d_Tt1o, const double* d_Tt2o,
                     const double* d_state, const double* d_a, const double* d_b,
                     const double* d_dc, const double* d_f0, const double* d_AdotV1,
                     const double* d_AdotV2, const double* d_dsigmadt,
                     const double* d_shear_loading, const double* d_slipv,
                     const double* d_cellpar, double* d_yTno, double* d_yTt1o, double* d_yTt2o,
                     double* d_ystate, unsigned long long* d_err, double a63, int64_t n_inst,
                     double* d_lstate, void* stream);

/* Commit the instances with d_flag 1 (as pq3d_rk45_commit, the state update
 * of src/QDsim.py:960-993), then t += h, h = dtnext, rejection count 0,
 * accepted-step count + 1, flag 0. Instances still attempting keep state, h
 * and count; they continue in the next pq3d_rk45_rounds of the same step.
 * Optional (NULL: unused, none has a reference counterpart; they replace
 * the driver loop's own bookkeeping, src/main.py:157-223):
 *   d_run     per-instance run record (PQ3D_RUN_W doubles): [0] end time
 *             (once t >= it the instance finishes: flag 3, no more attempts
 *             or commits), [1] armed, [2] event count, [3] peak slip rate,
 *             [4] event threshold up, [5] re-arm threshold down, [6]
 *             rejected attempts of the accepted steps, [8..] event start
 *             times (t after the first step whose max slip rate exceeds [4]
 *             while armed; re-armed below [5]);
 *   d_samples with d_lstate: one training row per accepted step
 *             (PQ3D_SAMPLE_W floats: the six features of the step's start,
 *             log10(h / dtnext base), 1) at row nacc*n_inst + b while it is
 *             below sample_cap;
 *   d_facc with d_b, d_f0 (per cell) and V0 > 0: the feature slots of
 *             pq3d_features (4 per instance); the accepted instances' slots
 *             are cleared and recomputed from their new state within the
 *             commit, so they equal what pq3d_features would give for those
 *             instances at the start of their next step (the slots of the
 *             other instances are left unchanged). */
int pq3d_rk45_advance(int64_t n, const int32_t* d_inst, double* d_h, uint8_t* d_flag,
                      int32_t* d_nrj, const double* d_dtnext, double* d_t, int64_t* d_nacc,
                      const double* d_a, const double* d_cellpar, const double* d_yTno,
                      const double* d_yTt1o, const double* d_yTt2o, const double* d_ystate,
                      double* d_Tno, double* d_Tt1o, double* d_Tt2o, double* d_state,
                      double* d_slipv1, double* d_slipv2, double* d_slipv, double* d_rake,
                      double* d_slip1, double* d_slip2, unsigned long long* d_vmax,
                      double* d_run, const double* d_lstate, float* d_samples,
                      int64_t sample_cap, int64_t n_inst, const double* d_b,
                      const double* d_f0, double V0, uint64_t* d_facc, void* stream);

/* Widths of the controller records, in this order: run record
 * (PQ3D_RUN_W = 64), learned-state record (PQ3D_LS_W = 16), training row
 * (PQ3D_SAMPLE_W = 8). Returns the number written (at most n). */
int pq3d_controller_widths(int64_t* out, int64_t n);

/* ------------------------------------------------- learned step size
 * 6 -> 64 -> 64 -> 64 -> 1 tanh MLP in FP32. Parameters packed as
 * W1 64x6 | b1 | W2 64x64 | b2 | W3 64x64 | b3 | W4 64 | b4 (row-major).
 * No reference counterpart (the reference uses only its controller,
 * src/QDsim.py:931-958); the proposal only seeds the first RK45 attempt. */

/* Number of packed parameters (8833). */
int pq3d_mlp_num_params(void);

/* Forward over nsamp inputs d_X (nsamp x 6); predictions to d_out when not
 * NULL; with d_grad != NULL also backward of 0.5 (y - T)^2: chunk samples per
 * workgroup, per-workgroup gradient sums to d_grad (ceil(nsamp/chunk) x P)
 * and loss sums to d_loss. */
int pq3d_mlp_forward_backward(const float* d_prm, const float* d_X, const float* d_T,
                              int64_t nsamp, int64_t chunk, float* d_out, float* d_grad,
                              float* d_loss, void* stream);

/* Sum the nblk gradient slots in order, divide by nsamp, write the gradient
 * to d_gout (optional) and, if lr > 0, take Adam step t on d_prm. */
int pq3d_mlp_adam(float* d_prm, const float* d_grad, int64_t nblk, int64_t nsamp, float* d_m,
                  float* d_v, float lr, float beta1, float beta2, float eps, int t, float* d_gout,
                  void* stream);

/* Per-instance solver features over ne cells, cell e of instance d_inst[e]
 * (NULL: one instance; 1 <= nb <= 1024), as 64-bit integers independent of summation
 * order: d_acc[4b] and d_acc[4b+1] = largest bit patterns of the doubles
 * 1000 + log10 V and 1000 - log10 V (log10 V clamped to [-999, 999]),
 * d_acc[4b+2] = sum of log10(V theta / Dc) in units of 2^-32 (two's
 * complement), d_acc[4b+3] = count(V > 1e-3); theta = Dc/V0 exp((psi-f0)/b).
 * No counterpart in the reference (it has no learned step size). */
int pq3d_features(int64_t ne, int nb, const int32_t* d_inst, const double* d_slipv,
                  const double* d_psi, const double* d_b, const double* d_f0, double V0,
                  uint64_t* d_acc, void* stream);

/* First-attempt step size of the device controller for the instances that
 * begin a fresh step (d_flag 0, d_nrj 0): the six features (d_acc of
 * pq3d_features over the current state divided by d_ncell[b] cells, and the
 * last accepted max(c1, c2) and dtnext/h from d_lstate) and the dtnext base
 * (d_h) go into d_lstate; with d_prm (NULL: record only) the network's
 * prediction p (inputs normalised by d_norm[0..11], output scaled by
 * d_norm[13] and shifted by d_norm[12]) sets d_h = base 10^min(p - d_norm[14],
 * d_norm[15]) and marks the step learned. No reference counterpart. */
int pq3d_learned_propose(int64_t n_inst, const uint64_t* d_acc, const int64_t* d_ncell,
                         const uint8_t* d_flag, const int32_t* d_nrj, double* d_h,
                         double* d_lstate, const float* d_prm, const float* d_norm,
                         void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PYQUAKE3D_AMD_H */
