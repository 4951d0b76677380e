/*
 * b2g: MI355X (gfx950) two-site DMRG effective-Hamiltonian hot path.
 *
 * C ABI of libb2g.so.  Plain pointers and sizes, int status codes
 * (B2G_OK == 0), no exceptions across the boundary, error text through
 * b2g_last_error().  All device memory is owned by a context.
 *
 * Two layers are exported:
 *
 *  (1) The thin device layer (b2g_ctx_*, b2g_buf_*, b2g_upload_plan,
 *      b2g_hpsi, b2g_diag, b2g_vec_*): a recorded list of GEMM-chain pairs
 *      (the device image of BatchGEMMSeq's recorded batches) replayed on the
 *      GPU.  It replaces the replay side of the reference:
 *        TensorFunctions::operator()(b, c, scale)      src/core/tensor_functions.hpp:59-62
 *        BatchGEMMSeq::operator()(c, v)                 src/core/batch_gemm.hpp:1570-1691
 *        AdvancedGEMM::tensor_product_diagonal          src/core/batch_gemm.hpp:507-513
 *
 *  (2) The effective-Hamiltonian layer (b2g_effham_*): takes the block-sparse
 *      problem exactly as block2's EffectiveHamiltonian constructor receives it
 *      (operator infos, operator blocks, the H_eff expression, bra/ket infos),
 *      builds the connection lists and the pair list on the host, uploads them
 *      and offers H*psi, the diagonal and the Davidson solve.  It replaces
 *        EffectiveHamiltonian<S,FL,MPS>  src/dmrg/effective_hamiltonian.hpp:99-1162
 *          ctor (diag + wfn_infos)        :126-224
 *          precompute / post_precompute   :226-253
 *          operator()(b, c, idx, factor)  :459-477
 *          eigs(...) -> (E, ndav, nflop, tdav)  :480-568
 *          deallocate()                   :1136-1162
 *        EffectiveHamiltonian<S,FL,MultiMPS>::eigs (k roots)  :1708-1800
 *        tensor_left/right_partial_expectation (backward)
 *                                         src/core/operator_functions.hpp:354-472
 */
#ifndef B2G_H
#define B2G_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------------- status */
#define B2G_OK 0
#define B2G_ERR_ARG 1     /* invalid argument / inconsistent shapes        */
#define B2G_ERR_DEVICE 2  /* HIP runtime error                             */
#define B2G_ERR_MEMORY 3  /* allocation failed                             */
#define B2G_ERR_NOCONV 4  /* Davidson hit max_iter before convergence      */
#define B2G_ERR_STATE 5   /* object used in the wrong state                */

/* element types */
#define B2G_F64 0
#define B2G_F32 1

/* precision of the H*psi contraction (b2g_hpsi / b2g_effham_*) */
#define B2G_PREC_F64 0    /* FP64 MFMA, FP64 accumulate                    */
#define B2G_PREC_F32 1    /* FP32 MFMA, FP32 accumulate                    */
#define B2G_PREC_BF16X3 2 /* split-BF16 three-term product, FP32 accumulate */

/* symmetry of the packed quantum labels (block2 SZLongLong / SU2LongLong
 * bit layouts, src/core/symmetry.hpp:656-732, 1184-1290)                  */
#define B2G_SZ 0
#define B2G_SU2 1

typedef struct b2g_ctx_s *b2g_ctx;
typedef struct b2g_buf_s *b2g_buf;
typedef struct b2g_plan_s *b2g_plan;
typedef struct b2g_effham_s *b2g_effham;

/* sizeof of the ABI structs, in this order: b2g_pair, b2g_block,
 * b2g_diag_term, b2g_sparse_info, b2g_operator, b2g_term, b2g_tp_conn,
 * b2g_problem, b2g_davidson_opts.  Bindings check their layouts with it.  Returns the count. */
int b2g_abi_sizes(size_t *out, int n);
/* Last error message of the calling thread ("" if none). */
const char *b2g_last_error(void);

/* ------------------------------------------------------------ context */
/* One context per host thread; it owns a HIP stream on `device`.        */
int b2g_ctx_create(int device, unsigned flags, b2g_ctx *out);
int b2g_ctx_destroy(b2g_ctx ctx);
int b2g_ctx_synchronize(b2g_ctx ctx);
/* the context's hipStream_t (for callers that order their own work or
 * record events against it)                                              */
void *b2g_ctx_stream(b2g_ctx ctx);
/* number of compute units of the context's device */
int b2g_ctx_num_cu(b2g_ctx ctx, int *out);

/* -------------------------------------------------------------- buffers */
/* Device buffer of n elements of dtype (B2G_F64 / B2G_F32).              */
int b2g_buf_create(b2g_ctx ctx, size_t n, int dtype, b2g_buf *out);
int b2g_buf_destroy(b2g_buf buf);
/* Device buffer initialised from host memory.  Replaces the per-site
 * upload of L/R operator blocks (the data behind lopt/ropt->ops).        */
int b2g_upload_blocks(b2g_ctx ctx, const void *host, size_t n, int dtype,
                      b2g_buf *out);
int b2g_buf_write(b2g_buf buf, const void *host, size_t offset, size_t n);
int b2g_buf_read(b2g_buf buf, void *host, size_t offset, size_t n);
/* raw device address (for callers that hand device pointers in directly) */
void *b2g_buf_device_ptr(b2g_buf buf);
size_t b2g_buf_size(b2g_buf buf);

/* ------------------------------------------------------------ pair list */
/* One GEMM chain  V[v] += alpha * op(A) * C[c] * op(B)^T  as recorded by
 * OperatorFunctions::tensor_product_multiply -> BatchGEMMSeq::rotate
 * (src/core/operator_functions.hpp:474-542, src/core/batch_gemm.hpp:893-902,
 * src/core/matrix_functions.hpp:998-1009).  All blocks are row-major.
 *   conj bit 0 set:   op(A) = A^T  (A stored c_m x v_m) else A (v_m x c_m)
 *   conj bit 1 set:   C*B         (B stored c_n x v_n) else C*B^T (v_n x c_n)
 * Offsets are element offsets into the L buffer (A), R buffer (B), the input
 * vector (C) and the output vector (V).                                  */
typedef struct {
    uint64_t a_off, b_off, c_off, v_off;
    uint32_t c_m, c_n, v_m, v_n;
    uint32_t conj;
    /* 0: two-tensor rotate on whole blocks; 1: a three-tensor rotate
     * (BatchGEMMSeq::three_rotate, src/core/batch_gemm.hpp:952-1022,
     * GMatrixFunctions::three_rotate, matrix_functions.hpp:1046-1112) whose
     * C and V are windows of their blocks: a range of rows (delayed left
     * operator) or of columns (delayed right operator)                     */
    uint32_t kind;
    /* row strides of the C and V windows (0: dense, = c_n / v_n)          */
    uint32_t c_ld, v_ld;
    double alpha;
} b2g_pair;

/* Output-vector block table: every bra block (offset, rows, cols), so that
 * blocks no pair touches are written too.                                */
typedef struct {
    uint64_t off;
    uint32_t m, n;
} b2g_block;

/* Diagonal contribution  D[d](r0 + i, c0 + j) += alpha * A(i,i) * B(j,j)
 * (AdvancedGEMM::tensor_product_diagonal, src/core/batch_gemm.hpp:507-513;
 * GMatrixFunctions::tensor_product_diagonal, matrix_functions.hpp:1204-1212;
 * the three-tensor form, batch_gemm.hpp:516-565, adds a diagonal window).
 * A is a square block (a_n x a_n) at a_off in the L buffer, B square
 * (b_n x b_n) at b_off in the R buffer; the a_n x b_n window at (r0, c0)
 * lies inside the output block at d_off.                                  */
typedef struct {
    uint64_t a_off, b_off, d_off;
    uint32_t a_n, b_n;
    uint32_t r0, c0;
    double alpha;
} b2g_diag_term;

/* Compile and upload a replay plan.  `lbuf`/`rbuf` hold the operator blocks
 * (FP64); x_len / y_len are the input/output vector lengths.  `flags` bit 0
 * or bit 1 also builds FP32 shadows of the operator buffers, which
 * B2G_PREC_F32 reads.  B2G_PREC_BF16X3 reads packed split-BF16 copies of the
 * operator blocks (hi + lo planes, 4 bytes per element, in the orientation
 * each pair uses), built on its first use or by b2g_plan_reserve.          */
int b2g_upload_plan(b2g_ctx ctx, b2g_buf lbuf, b2g_buf rbuf,
                    const b2g_pair *pairs, size_t n_pairs,
                    const b2g_block *yblocks, size_t n_yblocks, size_t x_len,
                    size_t y_len, unsigned flags, b2g_plan *out);
int b2g_plan_destroy(b2g_plan plan);
/* Re-derive the plan's reduced-precision shadows and packed planes after
 * the operator buffers' contents changed (same structure, new numbers).   */
int b2g_plan_refresh(b2g_plan plan);
/* Allocate what b2g_hpsi needs for n_rhs right-hand sides at `precision`
 * (scratch, and the packed operator planes of B2G_PREC_BF16X3), so that a
 * following b2g_hpsi allocates nothing (e.g. under stream capture).        */
int b2g_plan_reserve(b2g_plan plan, int n_rhs, int precision);
/* Reference-order work of one H*psi, in multiply-adds
 * (sum over pairs of c_m*c_n*v_n + v_m*c_m*v_n, the reference's nflop,
 * src/core/batch_gemm.hpp:307).  FLOPs = 2 * this.                        */
int b2g_plan_macs(b2g_plan plan, uint64_t *out);
/* Kernel ids of one H*psi for timing and accounting: the grouped GEMM of
 * phase 1 (work = C op(B)) in tile configuration c is B2G_K_PHASE1 + c,
 * phase 2 (V += alpha op(A) work) is B2G_K_PHASE2 + c (c = 3*i + j,
 * TM = 32 << i, TN = 32 << j), the split-K reduction B2G_K_SPLITK_REDUCE,
 * and the split-BF16 packing of the wavefunction B2G_K_PACK.              */
#define B2G_K_PHASE1 0
#define B2G_K_PHASE2 9
#define B2G_K_SPLITK_REDUCE 18
#define B2G_K_PACK 19
#define B2G_NKERNELS 20
/* Per-launch kernel timing with HIP events on the context stream: while
 * enabled, every b2g_hpsi runs its launches back to back on the context
 * stream (no overlap on the auxiliary streams) and brackets each with two
 * events;
 * b2g_plan_timing waits for them, writes per kernel id the summed duration
 * (ms) and the number of launches into ms[B2G_NKERNELS] and
 * n_launch[B2G_NKERNELS], then clears the record.                        */
int b2g_plan_set_timing(b2g_plan plan, int enable);
int b2g_plan_timing(b2g_plan plan, double *ms, uint64_t *n_launch);
/* Algorithmic FLOPs of one H*psi per right-hand side, per kernel id: 2 x the
 * multiply-adds of the kernel's tiles (phase 2 sums to exactly the
 * reference's second-GEMM count; phase 1 computes each distinct work matrix
 * once).  flops[B2G_NKERNELS].                                            */
int b2g_plan_kernel_flops(b2g_plan plan, double *flops);
/* The same for the launches of one precision: B2G_PREC_BF16X3 counts the
 * tiles of the packed split-BF16 plan (built here if it is not yet), whose
 * tiling differs; the others are b2g_plan_kernel_flops.                   */
int b2g_plan_kernel_flops_for(b2g_plan plan, int precision, double *flops);
/* Relative throughput (any unit) of the 9 workgroup tile shapes
 * TM x TN, TM, TN in {32, 64, 128}, index 3*i + j with TM = 32 << i,
 * TN = 32 << j.  The tiling of plans uploaded afterwards minimises the
 * estimated time with these rates (the autotuner measures them).  Process
 * wide.                                                                   */
int b2g_set_tile_rates(const double *rates, int n);
/* The same for the kernels of one precision: B2G_PREC_F64 and B2G_PREC_F32
 * share the table above; B2G_PREC_BF16X3 has its own, with which the packed
 * split-BF16 plan is tiled when it is built.                              */
int b2g_set_tile_rates_for(int precision, const double *rates, int n);
/* The rate table currently installed for the kernels of `precision`.     */
int b2g_get_tile_rates_for(int precision, double *rates, int n);
/* Number of kernel tasks per phase and scratch size, for reporting.      */
int b2g_plan_stats(b2g_plan plan, uint64_t *n_tasks1, uint64_t *n_tasks2,
                   uint64_t *scratch_elems);

/* y[r] = (accumulate ? y[r] : 0) + scale * H * x[r]  for r < n_rhs.
 * x and y are device arrays of n_rhs stacked vectors (x_len, y_len each),
 * FP64.  precision: B2G_PREC_*.  Enqueued on the context stream (with one
 * right-hand side the independent launches of a phase overlap on the
 * context's auxiliary streams, forked from and joined back to it).
 * Replaces TensorFunctions::operator()(b, c, scale).                     */
int b2g_hpsi(b2g_ctx ctx, b2g_plan plan, const double *x, double *y,
             int n_rhs, double scale, int accumulate, int precision);
/* b2g_hpsi that allocates nothing: the W windows and split-K partials live
 * in a fixed 64 MiB device array of the library (one per device, held until
 * the process ends) instead of plan-owned memory.  No counterpart in the
 * reference: for callers that cannot allocate (code under stream capture,
 * serving loops, harnesses that forbid hipMalloc).  FP64 and FP32 only
 * (split-BF16 reads packed planes the plan allocates); a product needing
 * more than 64 MiB of scratch is refused (B2G_ERR_ARG).  Every call on one
 * device shares that array: calls that overlap on one device (other
 * streams or contexts, unordered) race.  Same results as b2g_hpsi, bit for
 * bit.                                                                    */
int b2g_hpsi_noalloc(b2g_ctx ctx, b2g_plan plan, const double *x, double *y,
                     int n_rhs, double scale, int accumulate, int precision);

/* Partial expectations (the backward of E = sum_r <v_r| H |c_r>): the
 * derivative of E with respect to every element of the left and right
 * operator buffers, gl (lbuf length) and gr (rbuf length), FP64 device
 * arrays, overwritten (accumulate = 0) or added to.  v: n_rhs stacked bra
 * vectors (y_len each), c: n_rhs ket vectors (x_len each).  The
 * derivative with respect to c is H^T v (b2g_hpsi of a symmetric H).
 * Replaces the contraction behind
 *   OperatorFunctions::tensor_left_partial_expectation / _right_
 *                                  src/core/operator_functions.hpp:354-472
 *   TensorFunctions::tensor_product_expectation
 *                                  src/core/tensor_functions.hpp:1703-1878
 * (per block instead of per operator: a caller sums the blocks of one
 * operator, E_w = f_w <X_w, R_w>).  The gradient plan is built from the
 * uploaded pair list on the first call and kept with the plan.          */
int b2g_grad_ops(b2g_ctx ctx, b2g_plan plan, const double *v, const double *c,
                 int n_rhs, double *gl, double *gr, int accumulate);
/* algorithmic FLOPs of one right-hand side of b2g_grad_ops              */
int b2g_grad_flops(b2g_plan plan, double *flops);

/* Diagonal of H into d (device, y_len doubles, overwritten).            */
int b2g_diag(b2g_ctx ctx, b2g_buf lbuf, b2g_buf rbuf,
             const b2g_diag_term *terms, size_t n_terms,
             const b2g_block *dblocks, size_t n_dblocks, double *d,
             size_t d_len);

/* ------------------------------------------------- device vector kernels */
/* FP64 BLAS-1 style kernels used by the Davidson loop; n-length device
 * vectors, deterministic reductions (fixed block order).                 */
int b2g_vec_dot(b2g_ctx ctx, const double *x, const double *y, size_t n,
                double *out_host);
int b2g_vec_axpy(b2g_ctx ctx, double a, const double *x, double *y,
                 size_t n);
int b2g_vec_scale(b2g_ctx ctx, double a, double *x, size_t n);
int b2g_vec_copy(b2g_ctx ctx, const double *x, double *y, size_t n);

/* ------------------------------------------------ effective Hamiltonian */
/* Block-sparse matrix info (SparseMatrixInfo, sparse_matrix.hpp:47-877):
 * n blocks with composite quanta (sorted ascending), row/col dims and
 * element offsets of each block.                                          */
typedef struct {
    uint64_t delta_quantum;
    int32_t is_fermion;
    int32_t n;
    const uint64_t *quanta;
    const uint32_t *n_states_bra;
    const uint32_t *n_states_ket;
    const uint64_t *block_shifts;
} b2g_sparse_info;

/* One operator block set (an entry of lopt/ropt->ops): delta quantum,
 * scalar factor, the index of its info in left/right_op_infos and the
 * element offset of its data in the operator buffer.  flags bit 0: the
 * operator has structure only (a delayed operator of a DelayedOperatorTensor
 * whose blocks are never formed; data_off is ignored).                    */
typedef struct {
    uint64_t delta_quantum;
    double factor;
    int32_t info;
    int32_t flags;
    uint64_t data_off;
} b2g_operator;

/* One string of the H_eff OpSum (expr.hpp:503, 619, 811).
 * kind 0, OpProduct:  H += factor * op(L[a]) (x) op(R[b]) with block2's
 *   conj bits.
 * kind 1, OpSumProd (a delayed three-tensor string): the operator of the
 *   delayed side (L[a] or R[b], structure only) is the tensor product of
 *   sub-operators dl[da] and dr[db] of the delayed tensor, conjugated by the
 *   bits of dconj (OpSumProd::conjs), and one of them is 1 x 1
 *   (OperatorFunctions::three_tensor_product_multiply,
 *   src/core/operator_functions.hpp:543-671).                            */
typedef struct {
    int32_t a, b;
    uint32_t conj;
    uint32_t kind;
    double factor;
    int32_t da, db;
    uint32_t dconj;
    uint32_t pad;
} b2g_term;

/* A tensor-product connection list (SparseMatrixInfo::ConnectionInfo built
 * by initialize_tp, src/core/sparse_matrix.hpp:294-478) of a delayed
 * operator's info: classes (dconj, combined quantum) in `quanta` with
 * first entries `idx`, n[c] the first class of conj c, n[4] the class
 * count; entries (ia: dl block, ib: dr block, ic: fused block, stride: the
 * sub-block's row * fused-block columns + column, factor).                */
typedef struct {
    int32_t n[5];
    int32_t nc;
    const uint64_t *quanta;
    const uint32_t *idx;
    const uint32_t *ia, *ib, *ic;
    const uint64_t *stride;
    const double *factor;
} b2g_tp_conn;

typedef struct {
    int32_t symmetry;          /* B2G_SZ / B2G_SU2                         */
    int32_t n_left_infos, n_right_infos;
    int32_t n_left_ops, n_right_ops, n_terms;
    uint64_t opdq;             /* H's delta quantum (hop->q_label)         */
    uint64_t hop_left_vacuum;  /* left vacuum of the MPO                   */
    b2g_sparse_info ket, bra;  /* psi (ket) and sigma (bra) layouts        */
    const uint64_t *left_info_dq, *right_info_dq; /* sorted keys           */
    const b2g_sparse_info *left_infos, *right_infos;
    const b2g_operator *left_ops, *right_ops;
    const b2g_term *terms;
    const double *left_data;   /* host FP64 operator blocks                */
    const double *right_data;
    size_t left_len, right_len;
    /* Delayed (three-tensor) strings, DelayedOperatorTensor
     * (src/core/operator_tensor.hpp:209-): delayed_side 0 (none), 1 (the
     * left tensor is delayed) or 2 (the right one); the delayed tensor's
     * sub-operator sets dl (its lopt) and dr (its ropt) with their own
     * infos, whose data_off index the delayed side's data buffer; tp[i]
     * the tensor-product connection list of the delayed side's info i
     * (zeroed entries where none).                                        */
    int32_t delayed_side;
    int32_t n_dl_infos, n_dr_infos, n_dl_ops, n_dr_ops;
    int32_t pad_delayed;
    const b2g_sparse_info *dl_infos, *dr_infos;
    const b2g_operator *dl_ops, *dr_ops;
    const b2g_tp_conn *tp;
} b2g_problem;

/* Build connection lists, pair list, diagonal; upload.  flags as in
 * b2g_upload_plan; with B2G_EFFHAM_HOST_ONLY only the host-side plan is
 * built (connection lists and pairs can be queried, nothing touches a
 * device).                                                                 */
#define B2G_EFFHAM_HOST_ONLY 0x100
/* with B2G_EFFHAM_DEVICE_DATA, left_data / right_data are device addresses
 * (FP64, on the context's device); they are copied into the context.      */
#define B2G_EFFHAM_DEVICE_DATA 0x200
int b2g_effham_create(b2g_ctx ctx, const b2g_problem *prob, unsigned flags,
                      b2g_effham *out);
int b2g_effham_destroy(b2g_effham h);
/* sizes: |psi| (ket total), |sigma| (bra total), pairs, reference MACs    */
int b2g_effham_info(b2g_effham h, uint64_t *ket_len, uint64_t *bra_len,
                    uint64_t *n_pairs, uint64_t *macs);
/* The diagonal's terms: their number, the summed term windows (multiply-
 * adds of one diagonal) and the diagonal elements of the operator blocks
 * they read.  For accounting only.                                       */
int b2g_effham_diag_stats(b2g_effham h, uint64_t *n_terms,
                          uint64_t *window_elems, uint64_t *read_elems);
/* Host copies of the connection lists the builder produced (for checking
 * against ConnectionInfo): which = 0 wfn list, 1 diagonal list.  Arrays of
 * length *n_conn (queried with NULL pointers first).                      */
int b2g_effham_connections(b2g_effham h, int which, uint64_t *n_conn,
                           uint32_t *ia, uint32_t *ib, uint32_t *ic,
                           uint64_t *stride, double *factor);
int b2g_effham_pairs(b2g_effham h, uint64_t *n_pairs, b2g_pair *pairs);
/* the replay plan behind an effective Hamiltonian (NULL if host-only)    */
b2g_plan b2g_effham_plan(b2g_effham h);
/* H*psi on device arrays (see b2g_hpsi).                                  */
int b2g_effham_hpsi(b2g_effham h, const double *x, double *y, int n_rhs,
                    double scale, int accumulate, int precision);
/* Copy the diagonal (device, bra_len doubles).                            */
int b2g_effham_diag(b2g_effham h, double *d);
/* Replace the operator data (device FP64 arrays of the left / right
 * buffer lengths; same structure): copies them in, refreshes the
 * reduced-precision shadows and the diagonal.  For callers that own the
 * numbers (e.g. an autograd graph) while the plan stays.                 */
int b2g_effham_set_data(b2g_effham h, const double *left, const double *right);
/* The effective Hamiltonian's own device operator buffers and lengths.   */
int b2g_effham_data(b2g_effham h, double **left, size_t *left_len,
                    double **right, size_t *right_len);
/* Davidson driver of b2g_effham_eigs for one root: mode 1 (default)
 * captures one iteration (H*psi of the newest vector, subspace matrix,
 * Ritz pairs by device Jacobi, rotation, residual, Olsen, restart,
 * orthogonalisation, append) into a hipGraph and replays it, the host
 * reading a status header per iteration; mode 0 runs the host-driven loop
 * (also used for k > 1 roots and while plan timing is enabled).           */
int b2g_effham_set_davidson_mode(b2g_effham h, int mode);
/* User hook around every Davidson matvec (EffectiveKernel::compute,
 * src/dmrg/effective_hamiltonian.hpp:83-93, used at :522-532): with a hook
 * set, b2g_effham_eigs calls kernel(user, beta, f, f_data, a, b, n) where it
 * would compute b = beta * H a, and the hook calls f(f_data, a, b, scale)
 * (b = scale * H a on device vectors of length n) as it sees fit.  The
 * context stream is idle when the hook is entered and must be idle again
 * (its own work finished) when it returns; a non-zero return aborts the
 * solve with B2G_ERR_STATE.  A hook runs the host-driven Davidson loop.
 * kernel = NULL removes the hook.                                        */
typedef int (*b2g_matvec_fn)(void *f_data, const double *a, double *b,
                             double scale);
typedef int (*b2g_kernel_fn)(void *user, double beta, b2g_matvec_fn f,
                             void *f_data, const double *a, double *b,
                             size_t n);
int b2g_effham_set_kernel(b2g_effham h, b2g_kernel_fn kernel, void *user);
/* b2g_grad_ops on the effective Hamiltonian's plan.                      */
int b2g_effham_grad_ops(b2g_effham h, const double *v, const double *c,
                        int n_rhs, double *gl, double *gr, int accumulate);
/* Davidson (IterativeMatrixFunctions::davidson,
 * src/core/iterative_matrix_functions.hpp:866-1173, Olsen preconditioner):
 * k roots from k initial device vectors `x` (overwritten with the Ritz
 * vectors).  energies: host array of k.  Returns B2G_ERR_NOCONV when
 * max_iter is reached.  precision B2G_PREC_F64, or a mixed mode where the
 * matvecs run in FP32 / split-BF16 and a final FP64 stage refines
 * (refine != 0).                                                           */
int b2g_effham_eigs(b2g_effham h, double *x, int k, double conv_thrd,
                    int max_iter, int def_min, int def_max, int precision,
                    int refine, double *energies, int *ndav,
                    uint64_t *nflop, double *tdav);
/* DavidsonTypes (src/core/matrix_functions.hpp:281-301) that
 * b2g_effham_eigs_ex takes: the target order (Normal: lowest; GreaterThan,
 * LessThan, CloseTo a shift), the preconditioner (Olsen by default,
 * DavidsonPrecond, NoPrecond), Exact (dense diagonalisation) and
 * ElementProj.  Harmonic, NonHermitian and LeftEigen are refused.         */
#define B2G_DAV_NORMAL 0
#define B2G_DAV_GREATER_THAN 1
#define B2G_DAV_LESS_THAN 2
#define B2G_DAV_CLOSE_TO 4
#define B2G_DAV_DAVIDSON_PRECOND 32
#define B2G_DAV_NO_PRECOND 64
#define B2G_DAV_EXACT 256
#define B2G_DAV_ELEMENT_PROJ 1024
/* The arguments of EffectiveHamiltonian::eigs
 * (src/dmrg/effective_hamiltonian.hpp:480-490) beyond the vectors:
 * ortho holds n_ortho device vectors of ket length (ortho_bra: projected out
 * of the subspace and the residual, or, with proj_weights (host, n_ortho
 * values) and without B2G_DAV_ELEMENT_PROJ, added to H as w|v><v|;
 * iterative_matrix_functions.hpp:891-954, 974-976, 1071-1078).
 * Convergence: |r|^2 < conv_thrd + |lambda|^2 rel_conv_thrd^2.
 * soft_max_iter >= 0 stops after that many iterations without an error
 * (roots not converged report energy 0, as the reference).               */
typedef struct b2g_davidson_opts {
    double conv_thrd, rel_conv_thrd, shift;
    int max_iter, soft_max_iter;
    int def_min, def_max;
    int davidson_type;
    int precision, refine;
    int n_ortho;
    const double *const *ortho;
    const double *proj_weights;
} b2g_davidson_opts;
/* the reference's defaults: conv 5e-6, rel 0, max_iter 5000, no soft limit,
 * deflation 2 / 50, Normal, FP64                                           */
void b2g_davidson_opts_default(b2g_davidson_opts *o);
/* Davidson with the full option set (EffectiveHamiltonian::eigs /
 * EffectiveHamiltonian<S, FL, MultiMPS>::eigs, effective_hamiltonian.hpp:
 * 480-568, 1708-1800; IterativeMatrixFunctions::davidson,
 * iterative_matrix_functions.hpp:866-1173, exact_diagonalization :110-209):
 * k roots from the k device vectors `x`, overwritten with the targeted
 * Ritz vectors; energies (host, k).  Normal-type solves without orthogonal
 * states and hook run graph-captured on the device (b2g_effham_set_
 * davidson_mode 1, the default), everything else on the host-driven
 * loop.  ndav counts iterations, nflop the reference's multiply-adds of
 * all products.                                                            */
int b2g_effham_eigs_ex(b2g_effham h, double *x, int k,
                       const b2g_davidson_opts *opts, double *energies,
                       int *ndav, uint64_t *nflop, double *tdav);

/* ------------------------------------------------------------ truncation */
/* Batched symmetric eigensolver (GMatrixFunctions::eigs,
 * src/core/matrix_functions.hpp:1672-1688: xsyev on each block).  At most
 * 65535 matrices; matrix i of order n[i] (at most 16352, so that padded
 * blocks stay below 2^28 elements) is read row-major
 * from a + off[i] (device,
 * not modified).  Its eigenvalues go to w (device, compact: matrix i from
 * sum_{j<i} n[j]) in ascending order, its eigenvectors to v + off[i] as
 * rows (row j belongs to eigenvalue j), each signed so that its
 * largest-magnitude component is positive.  FP64 block Jacobi on the
 * device; sweeps (may be NULL) receives the number of Jacobi sweeps.     */
int b2g_syev_batched(b2g_ctx ctx, int n_mats, const uint32_t *n,
                     const uint64_t *off, const double *a, double *w,
                     double *v, int *sweeps);

/* Truncation types (TruncationTypes, src/dmrg/moving_environment.hpp:51-58):
 * selection by eigenvalue (PHYSICAL), eigenvalue / multiplicity (REDUCED)
 * or eigenvalue * multiplicity (REDUCED_INVERSED); SPECTRA_WITH_MULTIPLICITY
 * repeats each spectrum value by its multiplicity; KEEP_ONE * m keeps up to
 * m states of every sector beyond the bond dimension.                      */
#define B2G_TRUNC_PHYSICAL 0u
#define B2G_TRUNC_REDUCED 1u
#define B2G_TRUNC_REDUCED_INVERSED 2u
#define B2G_TRUNC_REAL_DENSITY_MATRIX 4u
#define B2G_TRUNC_SPECTRA_WITH_MULTIPLICITY 8u
#define B2G_TRUNC_KEEP_ONE 16u
/* block2's NoiseTypes values taken by b2g_split_density_matrix_noise */
#define B2G_NOISE_WAVEFUNCTION 1
#define B2G_NOISE_DENSITY_MATRIX 2

typedef struct b2g_split_s *b2g_split;
/* Density-matrix truncation of a two-site wavefunction on the device
 * (MovingEnvironment::density_matrix + split_density_matrix,
 * src/dmrg/moving_environment.hpp:3554-3577, 4250-4345; noise 0, one
 * wavefunction; trans_product scale = `scale`).  psi: device FP64 data of
 * the block-sparse wavefunction `wfn`.  trace_right != 0 (a forward sweep):
 * the result is the left rotation matrix (kept eigenvectors as columns)
 * and the right wavefunction L^T psi; trace_right == 0: the left
 * wavefunction psi R^T and the right rotation matrix (eigenvectors as
 * rows).  k: states to keep (-1: all), cutoff on the (reduced) eigenvalue,
 * normalize: 2-norm scaling of the wavefunction factor.  Eigenvectors are
 * signed as in b2g_syev_batched.                                           */
int b2g_split_density_matrix(b2g_ctx ctx, int symmetry,
                             const b2g_sparse_info *wfn, const double *psi,
                             int k, double cutoff, int trace_right,
                             int normalize, unsigned trunc_type, double scale,
                             b2g_split *out);
/* The split with perturbative noise (MovingEnvironment::density_matrix
 * with pkets, src/dmrg/moving_environment.hpp:3554-3577; the perturbed
 * wavefunctions of EffectiveHamiltonian::perturbative_noise,
 * src/dmrg/effective_hamiltonian.hpp:263-433, already scaled as
 * scale_perturbative_noise, :3698-3713, leaves them): the density matrix of
 * the wavefunction at offset 0 of `buf` (times `scale`) plus, for each of
 * the n_extra vectors of layout extra_infos[e] at buf + extra_offsets[e],
 * v v^T (trace_right) or v^T v on the density-matrix sectors it shares with
 * the wavefunction (OperatorFunctions::trans_product, operator_functions.hpp
 * :776-835, skips the others); then as b2g_split_density_matrix.         */
int b2g_split_density_matrix_extra(b2g_ctx ctx, int symmetry,
                                   const b2g_sparse_info *wfn,
                                   const double *buf, int n_extra,
                                   const b2g_sparse_info *extra_infos,
                                   const uint64_t *extra_offsets, int k,
                                   double cutoff, int trace_right,
                                   int normalize, unsigned trunc_type,
                                   double scale, b2g_split *out);
/* State-averaged form (MovingEnvironment::density_matrix_with_multi_target
 * + multi_split_density_matrix, src/dmrg/moving_environment.hpp:3582-3608,
 * 4347-4462): n_roots wavefunctions of the one structure `wfn`, stored one
 * after the other in psi (each of the structure's length), weights[r]
 * (host) scaling root r's density matrix (times `scale`).  One rotation
 * matrix; the wavefunction factor holds the n_roots new roots one after the
 * other (b2g_split_result's len is one root's), each normalised alone.    */
int b2g_split_density_matrix_multi(b2g_ctx ctx, int symmetry,
                                   const b2g_sparse_info *wfn,
                                   const double *psi, int n_roots,
                                   const double *weights, int k,
                                   double cutoff, int trace_right,
                                   int normalize, unsigned trunc_type,
                                   double scale, b2g_split *out);
int b2g_split_n_roots(b2g_split s, int *n);
/* block2's random generator as a handle (std::mt19937 seeded like
 * Random::rand_seed(seed), src/core/utils.hpp:196-252; seed 0, a clock
 * seed in block2, is refused).  Every draw continues its stream, as
 * block2's one generator continues across the sites of a sweep.          */
typedef struct b2g_rng_s *b2g_rng;
int b2g_rng_create(unsigned seed, b2g_rng *out);
int b2g_rng_destroy(b2g_rng rng);
/* With noise (MovingEnvironment::density_matrix with noise > 0 and
 * NoiseTypes Wavefunction or DensityMatrix, src/dmrg/moving_environment.hpp:
 * 3554-3577, through OperatorFunctions::trans_product,
 * src/core/operator_functions.hpp:776-835): the random tensor is drawn on
 * the host from std::mt19937 seeded with `seed`, in block2's order
 * (Random::fill, src/core/utils.hpp:226-252), so a reference run whose
 * generator was seeded with the same value (Random::rand_seed(seed))
 * adds the same noise.  Every call starts a new generator, while block2
 * draws from one stream that continues across sites: seed 0 (a clock seed
 * in block2) is refused with B2G_ERR_ARG when noise != 0, and a sweep
 * passes a new seed at each site.  Density-matrix noise leaves the
 * density matrix symmetric from its lower triangle, the part block2's
 * xsyev reads.  n_roots / weights as in b2g_split_density_matrix_multi
 * (density_matrix_with_multi_target, moving_environment.hpp:3582-3608:
 * each root draws its own tensor from the continuing stream); one root
 * with weight 1 is density_matrix's case.                                */
int b2g_split_density_matrix_noise(b2g_ctx ctx, int symmetry,
                                   const b2g_sparse_info *wfn,
                                   const double *psi, int n_roots,
                                   const double *weights, double noise,
                                   int noise_type, unsigned seed, int k,
                                   double cutoff, int trace_right,
                                   int normalize, unsigned trunc_type,
                                   double scale, b2g_split *out);
/* The same, drawing the noise from `rng` (b2g_rng_create): a sweep that
 * seeds one generator before its first site reproduces block2's noise at
 * every site.                                                            */
int b2g_split_density_matrix_noise_rng(b2g_ctx ctx, int symmetry,
                                       const b2g_sparse_info *wfn,
                                       const double *psi, int n_roots,
                                       const double *weights, double noise,
                                       int noise_type, b2g_rng rng, int k,
                                       double cutoff, int trace_right,
                                       int normalize, unsigned trunc_type,
                                       double scale, b2g_split *out);
/* Result parts: which 0 = left factor, 1 = right factor, 2 = the density
 * matrix (before diagonalisation).  The info arrays are owned by the
 * split; data is device memory owned by the split (the wavefunction
 * factor: n_roots consecutive roots of len elements each).                */
int b2g_split_result(b2g_split s, int which, b2g_sparse_info *info,
                     double **data, size_t *len);
/* Truncation error (sum of the discarded positive eigenvalues).          */
int b2g_split_error(b2g_split s, double *error);
/* which 0: all eigenvalues, per density-matrix block in ascending order;
 * which 1: the wavefunction spectra sqrt(max(0, w)) (store_wfn_spectra).
 * *n: capacity in, count out (out = NULL queries the count).             */
int b2g_split_values(b2g_split s, int which, double *out, size_t *n);
int b2g_split_destroy(b2g_split s);

/* ------------------------------------------------------------ file formats */
/* block2's binary sparse-matrix files (SparseMatrix::save_data /
 * load_data, src/core/sparse_matrix.hpp:908-989, with the info first when
 * saved with save_info: SparseMatrixInfo::save_data / load_data, :504-575),
 * read into and written from host memory.  ubond_size: bytes of block2's
 * ubond_t (4 in its large-bond build, else 2 or 1).  flags: the info's
 * last byte, block_shifts_fl_size << 2 | is_wavefunction (block2 writes
 * 8 << 2 | 1 for an FP64 wavefunction).  Compressed storage is refused.   */
typedef struct b2g_smat_s *b2g_smat;
int b2g_smat_load(const char *path, int with_info, int ubond_size,
                  b2g_smat *out);
/* info (requires with_info at load), flags, factor, data, len: any may be
 * NULL; arrays stay owned by the handle.                                  */
int b2g_smat_get(b2g_smat m, b2g_sparse_info *info, int *flags,
                 double *factor, const double **data, size_t *len);
/* info NULL: the matrix part only (save_data without save_info).          */
int b2g_smat_save(const char *path, const b2g_sparse_info *info, int flags,
                  int ubond_size, double factor, const double *data,
                  size_t len);
int b2g_smat_destroy(b2g_smat m);

/* ------------------------------- environment blocking and renormalisation */
/* The steps either side of H*psi in a sweep, on the device: forming the
 * operators of an enlarged block (left block (x) site, site (x) right
 * block), renormalising them with the MPS tensor, the normal-to-
 * complementary transform at the MPO's switch site, and the two-site
 * wavefunction moves between sites.
 *
 * A StateInfo (src/core/state_info.hpp:59-): n quanta in ascending order
 * with their state counts.                                                */
typedef struct {
    int32_t n;
    int32_t pad;
    const uint64_t *quanta;
    const uint32_t *n_states;
} b2g_state_info;

/* A window-sum plan: records  C_window += scale * op(A) (x) op(B)  grouped
 * by destination window, each window summed in record order by one
 * workgroup (no atomics).  It replays GMatrixFunctions::tensor_product
 * (src/core/matrix_functions.hpp:1269-) for block formation, and the block
 * copies of OperatorFunctions::iadd (src/core/operator_functions.hpp:
 * 134-174) and SparseMatrix::swap_to_fused_left / right
 * (src/core/sparse_matrix.hpp:1823-1968).                                 */
typedef struct b2g_winplan_s *b2g_winplan;

/* One string of a contraction expression: operator c of the result +=
 * factor * op(A[a]) (x) op(B[b]), conj bits as block2's (the OpProduct
 * strings of MPO::left_operator_exprs / right_operator_exprs entries,
 * divided by the result name's factor as TensorFunctions::left_contract
 * does).                                                                  */
typedef struct {
    int32_t c, a, b;
    uint32_t conj;
    double factor;
} b2g_contract_term;

/* TensorFunctions::left_contract / right_contract
 * (src/core/tensor_functions.hpp:2842-2986) through tensor_product
 * (:2184-2288) and OperatorFunctions::tensor_product
 * (src/core/operator_functions.hpp:672-711): the operators C of the
 * enlarged block on the fused bases bra = bra_a (x) bra_b, ket = ket_a (x)
 * ket_b (StateInfo::tensor_product, src/core/state_info.hpp:231-254; for a
 * left block A is the left environment and B the site, for a right block A
 * is the site and B the right environment).  Operator sets are given as
 * for an effective Hamiltonian: infos sorted by delta quantum, data
 * offsets into the A, B and C buffers.  The tensor-product connection lists of every C info
 * (ConnectionInfo::initialize_tp, src/core/sparse_matrix.hpp:294-422, for
 * the sub-labels of Partition::get_uniq_sub_labels,
 * src/dmrg/partition.hpp:368-480) are built here.                        */
typedef struct {
    int32_t symmetry;
    int32_t n_terms;
    b2g_state_info bra_a, ket_a, bra_b, ket_b, bra, ket;
    int32_t n_a_infos, n_b_infos, n_c_infos;
    int32_t n_a_ops, n_b_ops, n_c_ops;
    const uint64_t *a_info_dq, *b_info_dq, *c_info_dq;
    const b2g_sparse_info *a_infos, *b_infos, *c_infos;
    const b2g_operator *a_ops, *b_ops, *c_ops;
    const b2g_contract_term *terms;
} b2g_contract;

/* flags: B2G_EFFHAM_HOST_ONLY builds the host records only (queries work,
 * b2g_winplan_run refuses).                                               */
int b2g_contract_plan(b2g_ctx ctx, const b2g_contract *c, unsigned flags,
                      b2g_winplan *out);
/* The tensor-product connection list of C info `info` of a contraction
 * plan, in block2's ConnectionInfo layout (n[5]; classes: quanta, idx;
 * entries: ia, ib, ic, stride, factor).  Counts in *n_classes / *n_conn;
 * any array may be NULL.                                                   */
int b2g_winplan_connections(b2g_winplan w, int info, int32_t *n,
                            uint64_t *n_classes, uint64_t *quanta,
                            uint32_t *idx, uint64_t *n_conn, uint32_t *ia,
                            uint32_t *ib, uint32_t *ic, uint64_t *stride,
                            double *factor);

/* One string of a numerical transform: operator c += factor * op(a) (conj:
 * transposed, with block2's transpose factor), both in one operator set.  */
typedef struct {
    int32_t c, a;
    uint32_t conj;
    uint32_t pad;
    double factor;
} b2g_transform_term;

/* TensorFunctions::numerical_transform (src/core/tensor_functions.hpp:
 * 2462-2516) through OperatorFunctions::iadd: new operators of an
 * environment from the existing ones (the MPO's switch from normal to
 * complementary operators).  Run with a == c (one buffer): no term may read
 * an operator some term writes.                                           */
int b2g_transform_plan(b2g_ctx ctx, int symmetry, int n_infos,
                       const uint64_t *info_dq,
                       const b2g_sparse_info *infos, int n_ops,
                       const b2g_operator *ops, int n_terms,
                       const b2g_transform_term *terms, unsigned flags,
                       b2g_winplan *out);

/* SparseMatrix::swap_to_fused_left (to_left != 0: [l x (m r)] ->
 * [(l m) x r]) / swap_to_fused_right ([(l m) x r] -> [l x (m r)])
 * (src/core/sparse_matrix.hpp:1823-1968): old_fused / new_fused are the
 * fused bases (m (x) r and l (x) m, or the reverse) as MPSInfo::
 * swap_wfn_to_fused_left / right form them (src/dmrg/mps.hpp:716-790).   */
int b2g_swap_plan(b2g_ctx ctx, int symmetry, int to_left,
                  const b2g_state_info *l, const b2g_state_info *m,
                  const b2g_state_info *r, const b2g_state_info *old_fused,
                  const b2g_state_info *new_fused,
                  const b2g_sparse_info *old_wfn,
                  const b2g_sparse_info *new_wfn, unsigned flags,
                  b2g_winplan *out);

/* C = (accumulate ? C : 0) + the plan's sums.  a, b, c: device FP64.      */
int b2g_winplan_run(b2g_winplan w, const double *a, const double *b,
                    double *c, int accumulate);
/* windows, records, elements read (A and B) and written, and the length
 * of C the plan covers                                                    */
int b2g_winplan_stats(b2g_winplan w, uint64_t *n_groups, uint64_t *n_recs,
                      uint64_t *src_elems, uint64_t *dst_elems,
                      uint64_t *c_len);
int b2g_winplan_destroy(b2g_winplan w);

/* TensorFunctions::left_rotate / right_rotate (src/core/tensor_functions.hpp:
 * 2365-2403) through OperatorFunctions::tensor_rotate
 * (src/core/operator_functions.hpp:175-211): for every operator,
 * C[op] = rot^T A[op] rot (trans == 0, left) or rot A[op] rot^T (trans !=
 * 0, right), as a replay plan of GEMM-chain pairs (b2g_hpsi with x = the A
 * operator data, y = the C operator data, n_rhs 1).  rot: the MPS tensor
 * (delta quantum vacuum) whose data is rot_buf.  A and C operator sets as
 * in b2g_contract; a C operator is rotated from the A operator of the same
 * index.                                                                  */
int b2g_rotate_plan(b2g_ctx ctx, int symmetry, int trans,
                    const b2g_sparse_info *rot, b2g_buf rot_buf,
                    int n_a_infos, const uint64_t *a_info_dq,
                    const b2g_sparse_info *a_infos, int n_c_infos,
                    const uint64_t *c_info_dq,
                    const b2g_sparse_info *c_infos, int n_ops,
                    const b2g_operator *a_ops, const b2g_operator *c_ops,
                    size_t a_len, size_t c_len, unsigned flags,
                    b2g_plan *out);

/* A list of block products D[i] = scale * L[il] R[ir] on the device
 * (grouped-GEMM kernels): SparseMatrix::contract
 * (src/core/sparse_matrix.hpp:1777-1801), the two-site wavefunction of
 * MovingEnvironment::contract_two_dot (src/dmrg/moving_environment.hpp:
 * 3361-3404).  l_is_wfn != 0: wfn = wfn x rot (lmat a wavefunction, rmat
 * a right-canonical tensor); 0: wfn = rot x wfn.                          */
typedef struct b2g_gemmlist_s *b2g_gemmlist;
int b2g_contract_wfn_plan(b2g_ctx ctx, int symmetry,
                          const b2g_sparse_info *lmat, int l_is_wfn,
                          const b2g_sparse_info *rmat,
                          const b2g_sparse_info *out, double scale,
                          b2g_gemmlist *plan);
/* D = the products (every D block written; blocks without a product 0)  */
int b2g_gemmlist_run(b2g_gemmlist g, const double *l, const double *r,
                     double *d);
int b2g_gemmlist_stats(b2g_gemmlist g, uint64_t *n_products,
                       uint64_t *macs);
int b2g_gemmlist_destroy(b2g_gemmlist g);

/* A non-owning buffer over existing device memory (n elements of dtype):
 * destroying it leaves the memory alone.                                  */
int b2g_buf_wrap(b2g_ctx ctx, void *ptr, size_t n, int dtype, b2g_buf *out);

#ifdef __cplusplus
}
#endif

#endif /* B2G_H */
