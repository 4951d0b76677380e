// Device kernels of the H*psi replay, the diagonal and the Davidson vector
// algebra.  gfx950 (CDNA4) only: 64-lane wavefronts, FP64/FP32 MFMA.
#pragma once

#include <cstddef>

#include <hip/hip_runtime.h>

#include <cstdint>

namespace b2g {

// MFMA results land in their VGPRs several cycles after issue: VALU code that
// reads or overwrites accumulators must trail the last MFMA by at least 19
// wait states (16-pass FP64 MFMA; 8 for the BF16 16x16x32 one).  The
// compiler pads straight-line code, but not every path into a block that
// merges a k-loop's exits, so kernels call mfma_settle on their accumulators
// before such code: 24 wait states in one asm statement that takes every
// accumulator as an in/out operand, so that it follows all of their MFMAs
// and all later uses of them follow it.
#define B2G_SETTLE_NOPS "s_nop 7\n\ts_nop 7\n\ts_nop 7"
// (counts that are not a power of two are settled in power-of-two groups,
// each followed by its own wait states)
template <typename V, int N> __device__ __forceinline__ void mfma_settle(V *x) {
    if constexpr (N >= 16) {
        asm volatile(B2G_SETTLE_NOPS
                     : "+v"(x[0]), "+v"(x[1]), "+v"(x[2]), "+v"(x[3]),
                       "+v"(x[4]), "+v"(x[5]), "+v"(x[6]), "+v"(x[7]),
                       "+v"(x[8]), "+v"(x[9]), "+v"(x[10]), "+v"(x[11]),
                       "+v"(x[12]), "+v"(x[13]), "+v"(x[14]), "+v"(x[15]));
        mfma_settle<V, N - 16>(x + 16);
    } else if constexpr (N >= 8) {
        asm volatile(B2G_SETTLE_NOPS
                     : "+v"(x[0]), "+v"(x[1]), "+v"(x[2]), "+v"(x[3]),
                       "+v"(x[4]), "+v"(x[5]), "+v"(x[6]), "+v"(x[7]));
        mfma_settle<V, N - 8>(x + 8);
    } else if constexpr (N >= 4) {
        asm volatile(B2G_SETTLE_NOPS
                     : "+v"(x[0]), "+v"(x[1]), "+v"(x[2]), "+v"(x[3]));
        mfma_settle<V, N - 4>(x + 4);
    } else if constexpr (N >= 2) {
        asm volatile(B2G_SETTLE_NOPS : "+v"(x[0]), "+v"(x[1]));
        mfma_settle<V, N - 2>(x + 2);
    } else if constexpr (N == 1) {
        asm volatile(B2G_SETTLE_NOPS : "+v"(x[0]));
    }
}

// ---------------------------------------------------------------------------
// Segmented grouped GEMM.
//
// A *task* is one TM x TN tile of an output block D (row-major, ld = ldd).
// Its value is  D_tile = beta * D_tile + scale * sum_s alpha_s op(A_s) op(B_s)
// over the task's segments s, restricted to the tile.  Every output tile is
// owned by exactly one task, so accumulation is deterministic and needs no
// atomics.  Phase 1 of H*psi (work = C * op(B)) has one segment per task;
// phase 2 (V = sum alpha op(A) work) concatenates all pairs writing one bra
// block along K.
// ---------------------------------------------------------------------------
struct Seg {
    uint64_t a_off, b_off;
    uint32_t lda, ldb; // stored row lengths
    uint32_t K;
    uint32_t flags; // bit0: A transposed, bit1: B transposed
    double alpha;
};

struct Task {
    uint64_t d_off;
    uint32_t ldd, M, N; // block dims (for masking)
    uint32_t row0, col0;
    uint32_t seg_begin, seg_end;
    // bit 0: accumulate into D; bit 1: partial tile (split K) written to the
    // partial buffer P with d_off / ldd addressing the tile's dense slot
    uint32_t beta;
};

// Per-XCD task ranges of one launch: when off[8] != 0, workgroup b (dealt
// to XCD b % 8 by the dispatcher) walks tasks [off[b % 8], off[b % 8 + 1])
// with stride grid / 8; otherwise all workgroups stride the whole list.
struct XcdSplit {
    uint32_t off[9];
};

// Deterministic split-K reduction of one output tile:
//   D[tile] = (beta ? D[tile] : 0) + sum_{p < nparts} P[p_off + p*tm*tn + ...]
struct RedTask {
    uint64_t d_off, p_off;
    uint32_t ldd, M, N, row0, col0, tm, tn, nparts, beta, pad;
};

// One matrix of a packing job (split-BF16 operand planes, gemm_bf16x3.hip):
// a rows x dst_ld packed matrix at element offset dst (dst_ld = cols rounded
// up to 8), element (r, k) = src[src + r*src_ld + k] (trans:
// src[src + k*src_ld + r]) for k < cols and zero for cols <= k < dst_ld.  It
// is packed in groups of 8 consecutive k of one row, rows * dst_ld / 8 of
// them numbered from g0 on; the matrices of a job are sorted by g0.
struct PackMat {
    uint64_t src, dst;
    uint32_t src_ld, dst_ld;
    uint32_t rows, cols, trans, pad;
    uint64_t g0; // first 8-element group of the matrix in its pack job
};

// diagonal: output block with its contribution range
// rows [row0, row0 + m) of an output block of n columns, stored from d_off;
// terms [t_begin, t_end) are the block's terms whose rows meet the piece
struct DiagBlock {
    uint64_t d_off;
    uint32_t m, n;
    uint32_t t_begin, t_end;
    uint32_t row0, pad;
};
struct DiagTermDev {
    uint64_t a_off, b_off;
    uint32_t a_n, b_n;
    uint32_t r0, c0; // window origin inside the output block
    double alpha;
};

// Device-resident state of the graph-captured Davidson (davidson_dev.hip).
// The first 64 bytes are the header the host reads after each iteration.
struct DavState {
    // the first 128 bytes are the header the last kernel of a step copies
    // to pinned host memory
    int m;        // vectors in the subspace
    int status;   // 0 running, 1 converged, 2 iteration limit, 3 soft limit
    int iter;     // iterations done
    int max_iter;
    int def_min, def_max;
    int k;        // roots
    int ck;       // roots converged so far
    int appended; // the step appended a vector (the next product is needed)
    int soft_max_iter; // -1: none
    int type;     // DavidsonTypes preconditioner bits (32, 64)
    int pad0[5];
    double conv;   // threshold on the squared residual norm
    double qq;     // squared residual norm of the last iteration
    double energy; // lowest Ritz value of the last iteration
    double rel2;   // rel_conv_thrd^2
    double pad1[4];
    double scal[8]; // reduction results: |q|^2, Olsen c.q and c.t, |q|^2
    double ld[64], dots[64], rn[64];
    double G[64 * 64], V[64 * 64];
};
static_assert(offsetof(DavState, scal) == 128, "DavState header");


} // namespace b2g
