/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C CPU restatement of block2's two-site effective-Hamiltonian hot path,
 * used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
 * the checker.  The product (libb2g.so) never links or calls it.
 *
 * Restated from (file:line of the reference, block2-preview 0.5.4):
 *   labels        src/core/symmetry.hpp:656-732 (SZLongLong),
 *                 1184-1290 (SU2LongLong)
 *   couplings     src/core/clebsch_gordan.hpp:58-200 (wigner_6j, wigner_9j,
 *                 transpose_cg)
 *   sub-labels    src/dmrg/partition.hpp:368-480 (get_uniq_sub_labels)
 *   connections   src/core/sparse_matrix.hpp:81-161 (initialize_diag),
 *                 163-292 (initialize_wfn)
 *   H*psi         src/core/operator_functions.hpp:474-542
 *                 (tensor_product_multiply) + matrix_functions.hpp:998-1009
 *                 (rotate: work = C op(B)^T, V += alpha op(A) work)
 *   diagonal      src/core/operator_functions.hpp:212-247 +
 *                 matrix_functions.hpp:1204-1212
 *   Davidson      src/core/iterative_matrix_functions.hpp:93-108 (Olsen),
 *                 66-72 (Davidson), 866-1173 (davidson: Normal,
 *                 GreaterThan, LessThan, CloseTo, preconditioners, ors,
 *                 projection weights, rel_conv_thrd, soft_max_iter)
 *   partial expectations (gradients)
 *                 src/core/operator_functions.hpp:354-472
 *
 * Pinning: see tests/test_oracle.py (reference golden vectors from
 * oracle/_ref dumps, closed forms, brute force on tiny inputs).
 */
#ifndef B2G_ORACLE_H
#define B2G_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_problem orc_problem;

/* Flattened block-sparse infos: info i has n[i] blocks starting at off[i] in
 * the q / nb / nk / sh arrays. */
orc_problem *orc_create(
    int sym, uint64_t opdq, uint64_t hop_left_vacuum,
    /* ket */ uint64_t ket_dq, int ket_n, const uint64_t *ket_q,
    const uint32_t *ket_nb, const uint32_t *ket_nk, const uint64_t *ket_sh,
    /* bra */ uint64_t bra_dq, int bra_n, const uint64_t *bra_q,
    const uint32_t *bra_nb, const uint32_t *bra_nk, const uint64_t *bra_sh,
    /* left infos */ int nli, const uint64_t *li_dq, const uint32_t *li_ferm,
    const uint32_t *li_n, const uint32_t *li_off, const uint64_t *li_q,
    const uint32_t *li_nb, const uint32_t *li_nk, const uint64_t *li_sh,
    /* right infos */ int nri, const uint64_t *ri_dq, const uint32_t *ri_ferm,
    const uint32_t *ri_n, const uint32_t *ri_off, const uint64_t *ri_q,
    const uint32_t *ri_nb, const uint32_t *ri_nk, const uint64_t *ri_sh,
    /* left ops */ int nlo, const double *lo_f, const int32_t *lo_info,
    const uint64_t *lo_off, const double *ldata,
    /* right ops */ int nro, const double *ro_f, const int32_t *ro_info,
    const uint64_t *ro_off, const double *rdata,
    /* terms */ int nt, const uint8_t *t_conj, const int32_t *t_a,
    const int32_t *t_b, const double *t_f);
void orc_destroy(orc_problem *p);
const char *orc_error(orc_problem *p);

uint64_t orc_ket_len(const orc_problem *p);
uint64_t orc_bra_len(const orc_problem *p);
/* reference-order multiply-adds of one H*psi */
uint64_t orc_macs(const orc_problem *p);

/* which = 0: wavefunction list, 1: diagonal list.  Pass NULL arrays to
 * query the count. */
int64_t orc_connections(const orc_problem *p, int which, uint32_t *ia,
                        uint32_t *ib, uint32_t *ic, uint64_t *stride,
                        double *factor, int32_t *n5, int64_t *nq,
                        uint64_t *quanta, uint32_t *idx);

/* y = H x (y overwritten).  nthreads <= 0: OpenMP default. */
void orc_hpsi(const orc_problem *p, const double *x, double *y,
              int nthreads);
/* Same, restricted to the first max_pairs recorded pairs (bounded timing
 * sample); returns the multiply-adds performed. */
uint64_t orc_hpsi_sample(const orc_problem *p, const double *x, double *y,
                         uint64_t max_pairs, int nthreads);
/* y = H x restricted to the output (bra) blocks listed in `blocks` (the
 * other blocks of y are zero); each pair's GEMM rows are spread over
 * nthreads, every element summed in record order.  Returns the
 * multiply-adds performed. */
uint64_t orc_hpsi_blocks(const orc_problem *p, const double *x, double *y,
                         int nsel, const int32_t *blocks, int nthreads);
/* multiply-adds of one H*psi per output (bra) block */
void orc_block_macs(const orc_problem *p, uint64_t *macs);
void orc_diag(const orc_problem *p, double *d);
/* k roots; x holds k vectors (overwritten with the targeted Ritz vectors).
 * type: DavidsonTypes bits (Normal 0, GreaterThan 1, LessThan 2, CloseTo 4,
 * DavidsonPrecond 32, NoPrecond 64, ElementProj 1024); ors: nor vectors of
 * ket length (projected out, or added to H with proj_weights when that is not
 * NULL).  soft_max_iter -1: none.  nthreads: threads of the products.  Returns 0 on convergence, 1 if max_iter
 * was hit, 2 if soft_max_iter stopped it. */
int orc_davidson(const orc_problem *p, double *x, int k, double conv_thrd,
                 double rel_conv_thrd, int max_iter, int soft_max_iter,
                 int def_min, int def_max, int type, double shift, int nor,
                 const double *ors, const double *proj_weights,
                 double *energies, int *ndav, int nthreads);
/* Gradients of E = <v| H |c>:  gl / gr receive dE/d(left / right operator
 * data), laid out like the operator buffers (overwritten). */
void orc_grad_ops(const orc_problem *p, const double *v, const double *c,
                  double *gl, double *gr);

/* SU(2) symbols (doubled spins) for pinning against brute force. */
double orc_sixj(int a, int b, int c, int d, int e, int f);
double orc_ninej(int a, int b, int c, int d, int e, int f, int g, int h,
                 int i);

#ifdef __cplusplus
}
#endif

#endif
