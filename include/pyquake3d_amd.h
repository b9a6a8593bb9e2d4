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
 * ctypes at src/Hmatrix.py:25-38): full-space stress and strain at n receivers
 * (X[i],Y[i],Z[i]) from one triangle (P1,P2,P3, 3 doubles each) with Burgers
 * vector (Ss, Ds, Ts). Outputs n x 6 row-major. The half-space branch
 * (jud_halfspace = true) is not provided: outputs become NaN and
 * pq3d_last_error() says so. */
void TDstressFS_C(const double* X, const double* Y, const double* Z, size_t n,
                  const double* P1, const double* P2, const double* P3, double Ss, double Ds,
                  double Ts, double mu, double lambda_, bool jud_halfspace, double* stress,
                  double* strain);

/* Replaces `TDstressEachSourceAtReceiver_C` (reference
 * src/TDstressFS_C.cpp:2098-2143, bound at src/Hmatrix.py:40-49): one receiver
 * (x,y,z) and n_sources triangles given as flat n_sources x 3 arrays. Outputs
 * n_sources x 6 row-major. Same half-space behaviour as above. */
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

#ifdef __cplusplus
}
#endif

#endif /* PYQUAKE3D_AMD_H */
