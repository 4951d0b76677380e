// Device kernels of the H*psi replay, the diagonal and the Davidson vector
// algebra.  gfx950 (CDNA4) only: 64-lane wavefronts, FP64/FP32 MFMA.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace b2g {

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

// Deterministic split-K reduction of one output tile:
//   D[tile] = (beta ? D[tile] : 0) + sum_{p < nparts} P[p_off + p*tm*tn + ...]
struct RedTask {
    uint64_t d_off, p_off;
    uint32_t ldd, M, N, row0, col0, tm, tn, nparts, beta, pad;
};

// diagonal: output block with its contribution range
struct DiagBlock {
    uint64_t d_off;
    uint32_t m, n;
    uint32_t t_begin, t_end;
};
struct DiagTermDev {
    uint64_t a_off, b_off;
    uint32_t a_n, b_n;
    double alpha;
};

} // namespace b2g
