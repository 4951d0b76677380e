// Host-callable launchers of the device kernels (all enqueue on `st`).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace b2g {

// workgroup tile configuration cfg = 3 * i + j: TM = 32 << i, TN = 32 << j
int tile_m(int cfg);
int tile_n(int cfg);
constexpr int kTileConfigs = 9;

// persistent grid size of a grouped-GEMM launch (a multiple of 8 when the
// launch has per-XCD task ranges)
unsigned task_grid(int grid, int ntasks, const XcdSplit &xs);

// PHASE (1: work = C op(B), 2: V += alpha op(A) work, 3 / 4: the two stages
// of the operator gradients) only tags the kernel name, so that profiles tell
// the phases apart
template <typename TA, typename TB, typename TD, typename TC, int PHASE>
hipError_t launch_seg_gemm(int cfg, const TA *A, const TB *B, TD *D, TD *P,
                           const Task *tasks, int ntasks, const Seg *segs,
                           uint64_t a_rs, uint64_t b_rs, uint64_t d_rs,
                           uint64_t p_rs, double scale, int beta0, int n_rhs,
                           int grid, hipStream_t st, XcdSplit xs = XcdSplit{});
template <typename TD>
hipError_t launch_splitk_reduce(TD *D, const TD *P, const RedTask *tasks,
                                int ntasks, uint64_t d_rs, uint64_t p_rs,
                                int beta0, int n_rhs, int grid,
                                hipStream_t st);

// split-BF16 three-term variant on packed operand planes (gemm_bf16x3.hip):
// segments address packed A (M x lda) and packed B (N x ldb) planes, both
// k-contiguous; OUT 0 writes the packed transpose of the product as hi / lo
// planes (D0 / D1), OUT 1 writes FP64 D0 with split-K partials in D1
template <int OUT>
hipError_t launch_seg_gemm_pk(int cfg, const __bf16 *Ah, const __bf16 *Al,
                              const __bf16 *Bh, const __bf16 *Bl, void *D0,
                              void *D1, const Task *tasks, int ntasks,
                              const Seg *segs, uint64_t a_rs, uint64_t b_rs,
                              uint64_t d_rs, uint64_t p_rs, double scale,
                              int beta0, int n_rhs, int grid, hipStream_t st,
                              XcdSplit xs = XcdSplit{});
// FP64 -> packed hi / lo BF16 planes, per right-hand side (src_rs, dst_rs):
// the nmats matrices of `mats` (ngroups 8-element groups in all)
hipError_t launch_pack_bf16(const double *src, __bf16 *hi, __bf16 *lo,
                            const PackMat *mats, int nmats, uint64_t ngroups,
                            uint64_t src_rs, uint64_t dst_rs, int n_rhs,
                            int grid, hipStream_t st);

hipError_t launch_diag(const double *L, const double *R, double *Dv,
                       const DiagBlock *blocks, int nblocks,
                       const DiagTermDev *terms, int grid, hipStream_t st);
hipError_t launch_f64_to_f32(const double *s, float *t, size_t n,
                             hipStream_t st);
hipError_t launch_dot(const double *x, const double *y, size_t n,
                      double *partial, double *out, hipStream_t st);
hipError_t launch_gram(const double *X, int kx, const double *Y, int ky,
                       size_t ld, size_t n, double *partial, double *out,
                       hipStream_t st);
hipError_t launch_axpy(double a, const double *x, double *y, size_t n,
                       hipStream_t st);
hipError_t launch_axpy_dev(const double *coef, double sign, const double *x,
                           double *y, size_t n, hipStream_t st);
hipError_t launch_scale(double a, double *x, size_t n, hipStream_t st);
hipError_t launch_copy(const double *x, double *y, size_t n, hipStream_t st);
hipError_t launch_fill(double v, double *x, size_t n, hipStream_t st);
hipError_t launch_rotate(double *X, size_t ld, size_t n, int m,
                         const double *V, hipStream_t st);
hipError_t launch_olsen(double *q, const double *c, const double *aa,
                        double ld, double *t, size_t n, double *partial,
                        double *sums, hipStream_t st);
hipError_t launch_dav_precond(double *q, const double *aa, double ld,
                              size_t n, hipStream_t st);
int red_grid();

// One device Davidson step after sn = H qn (davidson_dev.hip): nv is the
// subspace capacity (2..64 vectors of B and S), `partial` holds
// davidson_partial_len() doubles, and the step ends by writing the state
// header (64 bytes) to `host` (pinned memory; nullptr: not written).
hipError_t launch_davidson_step(DavState *st, double *B, double *S,
                                const double *sn, double *q, double *t,
                                double *qn, const double *diag, size_t n,
                                int nv, double *partial, DavState *host,
                                hipStream_t s);
int davidson_partial_len();

} // namespace b2g
