// FP64 truncation of a two-site wavefunction on the device: the density
// matrix, its blockwise eigendecomposition, the selection of the kept states
// and the split into a rotation matrix and a wavefunction
//   MovingEnvironment::density_matrix          (noise 0, one wavefunction)
//       src/dmrg/moving_environment.hpp:3554-3577
//   OperatorFunctions::trans_product           src/core/operator_functions.hpp:776-835
//   SparseMatrixInfo::initialize_dm            src/core/sparse_matrix.hpp:677-724
//   MovingEnvironment::truncate_density_matrix src/dmrg/moving_environment.hpp:3716-3853
//   MovingEnvironment::split_density_matrix    src/dmrg/moving_environment.hpp:4250-4345
//   rotation_matrix_info_from_density_matrix / wavefunction_info_from_density_matrix
//                                              src/dmrg/moving_environment.hpp:4013-4083
//   GMatrixFunctions::eigs (xsyev per block)   src/core/matrix_functions.hpp:1672-1688
//
// Device work:
//   * density matrix blocks D[b] = sum over wavefunction blocks a of
//     psi_a psi_a^T (trace_right) or psi_a^T psi_a, as output-stationary
//     grouped-GEMM tasks (the H*psi engine, FP64 MFMA);
//   * a batched symmetric eigensolver: block Jacobi over 16-column blocks.
//     Each round pairs the column blocks of every matrix (round-robin); the
//     32 x 32 subproblem of a pair is diagonalised by cyclic Jacobi in LDS,
//     giving an orthogonal Q, and the whole matrix is updated as
//     A <- J^T A J, V <- V J with J the block-diagonal of the round's Qs:
//     every 32 x 32 region (pair P rows, pair R columns) becomes
//     Q_P^T A_PR Q_R.  Matrices are padded to a multiple of 16 with zero rows
//     and columns, which no rotation touches (a zero off-diagonal entry is
//     never rotated), so positions >= n stay out of the spectrum.  A global
//     sweep is max(nb) - 1 rounds (smaller matrices cycle through their own
//     rounds); the solve ends after a sweep with no rotation anywhere.
//     Off-diagonal entries below 64 eps ||A_b||_F are left (the rounding
//     level of a region update), so eigenvalues carry an absolute error of
//     that order, like xsyev's;
//   * eigenvectors gathered as rows in ascending eigenvalue order, each with
//     its largest-magnitude component positive (LAPACK leaves the sign open);
//   * the kept eigenvectors copied into the rotation matrix, and the other
//     factor (L^T psi, or psi R^T) as grouped-GEMM tasks; 2-norm scaling.
// Host work: block structure, the selection (sort, cutoff, error, KeepOne)
// in the reference's order, output infos.
#include "common.hpp"
#include "launch.hpp"
#include "replay.hpp"
#include "symmetry.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <random>
#include <string>
#include <utility>

namespace b2g {

namespace {

constexpr int kJB = 16;      // column block of the block Jacobi
constexpr int kJS = 2 * kJB; // order of a pair subproblem
// largest matrix order of the eigensolver: the padded order (a multiple of
// 2 kJB) squared stays below 2^28 elements
constexpr int kMaxSyevOrder = 16352;
static_assert((uint64_t)kMaxSyevOrder * kMaxSyevOrder < ((uint64_t)1 << 28) &&
                  kMaxSyevOrder % (2 * kJB) == 0,
              "kMaxSyevOrder");
constexpr int kJLD = kJS + 1;

// one matrix of a batch: A and V are N x N (N = 32 nb) from a_off; the
// round's Qs are nb / 2 matrices of 64 x 64 from q_off, their flags from
// f_off
struct JacBlock {
    uint64_t a_off, q_off;
    uint32_t nb, n, f_off, pad;
    double floor_abs; // off-diagonal entries below this are not rotated
};
// pair task (kind 0): x = pair index; region task (kind 1): x = row pair,
// y = column pair; V task (kind 2): x = pair, y = slice of kJS rows
struct JacTask {
    uint32_t blk, kind, x, y;
};

// Round-robin (circle) pairing of nb items (nb even): round r in [0, nb-1),
// pair k in [0, nb/2); every item meets every other once per nb - 1 rounds.
__device__ __forceinline__ void rr_pair(int nb, int r, int k, int &i, int &j) {
    const int m = nb - 1;
    int a, b;
    if (k == 0) {
        a = 0;
        b = r % m + 1;
    } else {
        a = (r + k) % m + 1;
        b = (r - k + m) % m + 1;
    }
    i = a < b ? a : b;
    j = a < b ? b : a;
}

// global row (or column) of local index l of the pair (bi, bj)
__device__ __forceinline__ int pair_index(int bi, int bj, int l) {
    return l < kJB ? bi * kJB + l : bj * kJB + (l - kJB);
}

// A: padded copy of the compact symmetric matrices, V = I
__global__ __launch_bounds__(256) void jacobi_init_kernel(
    const JacBlock *__restrict__ blocks, int nblocks,
    const uint64_t *__restrict__ src_off, const double *__restrict__ src,
    double *__restrict__ A, double *__restrict__ V) {
    for (int b = blockIdx.y; b < nblocks; b += gridDim.y) {
        const JacBlock bl = blocks[b];
        const uint32_t N = bl.nb * kJB;
        const uint64_t NN = (uint64_t)N * N;
        for (uint64_t e = blockIdx.x * 256ull + threadIdx.x; e < NN;
             e += (uint64_t)gridDim.x * 256) {
            const uint32_t i = (uint32_t)(e / N), j = (uint32_t)(e % N);
            A[bl.a_off + e] =
                (i < bl.n && j < bl.n) ? src[src_off[b] + (uint64_t)i * bl.n + j]
                                       : 0.0;
            V[bl.a_off + e] = i == j ? 1.0 : 0.0;
        }
    }
}

// floor_abs of each matrix: 64 eps ||A_b||_F.  Stage 1: workgroup
// (chunk, b) sums the squares of rows chunk, chunk + gridDim.x, ... of
// matrix b into part[b * gridDim.x + chunk]; stage 2 (jacobi_floor_kernel)
// adds the parts of each matrix in order (deterministic).
__global__ __launch_bounds__(256) void jacobi_norm_kernel(
    const JacBlock *__restrict__ blocks, const double *__restrict__ A,
    double *__restrict__ part) {
    __shared__ double sh[256];
    const JacBlock bl = blocks[blockIdx.y];
    const uint32_t N = bl.nb * kJB;
    double s = 0.0;
    for (uint32_t i = blockIdx.x; i < N; i += gridDim.x)
        for (uint32_t j = threadIdx.x; j < N; j += 256) {
            const double v = A[bl.a_off + (uint64_t)i * N + j];
            s = fma(v, v, s);
        }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = sh[0];
}

__global__ void jacobi_floor_kernel(JacBlock *__restrict__ blocks, int nblocks,
                                    const double *__restrict__ part,
                                    int nparts) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblocks)
        return;
    double s = 0.0;
    for (int c = 0; c < nparts; c++)
        s += part[(size_t)b * nparts + c];
    blocks[b].floor_abs = 64.0 * 2.220446049250313e-16 * sqrt(s);
}

// One kJS x kJS pair subproblem per workgroup: cyclic Jacobi in LDS,
// sweeps until one applies no rotation (at most max_sweeps); writes the
// accumulated rotation Q and whether it differs from I, and counts rotating
// pairs.  The kJB rotations of a round act on disjoint row / column pairs,
// so thread (k, l) applies rotation k to the rows and rotation l to the
// columns of the 2 x 2 block where they meet, reading its four entries
// before writing them; the diagonal block of each rotating pair gets its
// exact rotated values.  Two barriers per round: after the angles and
// after the block updates.
__global__ __launch_bounds__(256) void jacobi_pair_kernel(
    const JacBlock *__restrict__ blocks, const JacTask *__restrict__ tasks,
    const double *__restrict__ A, double *__restrict__ Q,
    int *__restrict__ qflag, int round, double tol, int max_sweeps,
    unsigned *__restrict__ nrot) {
    static_assert(kJB * kJB == 256, "one thread per 2 x 2 block of a round");
    __shared__ double S[kJS * kJLD];
    __shared__ double U[kJS * kJLD];
    __shared__ double rc[kJB], rs[kJB], rpp[kJB], rqq[kJB];
    __shared__ int rp[kJB], rq[kJB], rot[kJB];
    __shared__ int any_rot, need;
    __shared__ unsigned rmask[2]; // rotating pairs of a round, by parity
    const JacTask tk = tasks[blockIdx.x];
    const JacBlock bl = blocks[tk.blk];
    const uint32_t N = bl.nb * kJB;
    const int tid = threadIdx.x;
    // rotation test of entry (p, q): |a_pq| above the floor and above
    // tol sqrt(|a_pp a_qq|) (compared squared)
    const double tol2 = tol * tol;
    auto wants = [&](double apq, double app, double aqq) {
        return fabs(apq) > bl.floor_abs && apq * apq > tol2 * fabs(app * aqq);
    };
    int bi, bj;
    rr_pair((int)bl.nb, round % (int)(bl.nb - 1), (int)tk.x, bi, bj);
    for (int e = tid; e < kJS * kJS; e += 256) {
        const int l = e / kJS, c = e % kJS;
        S[l * kJLD + c] = A[bl.a_off + (uint64_t)pair_index(bi, bj, l) * N +
                            pair_index(bi, bj, c)];
        U[l * kJLD + c] = l == c ? 1.0 : 0.0;
    }
    int rotated = 0;
    for (int sw = 0; sw < max_sweeps; sw++) {
        if (tid == 0)
            any_rot = 0, need = 0;
        __syncthreads();
        // skip the sweep when no off-diagonal entry passes the test
        for (int e = tid; e < kJS * kJS; e += 256) {
            const int i = e / kJS, j = e % kJS;
            if (i < j && wants(S[i * kJLD + j], S[i * kJLD + i],
                               S[j * kJLD + j]))
                need = 1;
        }
        __syncthreads();
        if (!need)
            break;
        for (int r = 0; r < kJS - 1; r++) {
            int f = 0;
            if (tid < kJB) {
                int p, q;
                rr_pair(kJS, r, tid, p, q);
                const double apq = S[p * kJLD + q], app = S[p * kJLD + p],
                             aqq = S[q * kJLD + q];
                double c = 1.0, s = 0.0;
                if (wants(apq, app, aqq)) {
                    // symmetric Schur rotation zeroing a_pq
                    const double th = (aqq - app) / (2.0 * apq);
                    double t;
                    if (fabs(th) > 1e150)
                        t = 0.5 / th;
                    else
                        t = (th >= 0.0 ? 1.0 : -1.0) /
                            (fabs(th) + sqrt(th * th + 1.0));
                    c = rsqrt(t * t + 1.0);
                    s = t * c;
                    rpp[tid] = app - t * apq;
                    rqq[tid] = aqq + t * apq;
                    f = 1;
                    any_rot = 1;
                }
                rc[tid] = c;
                rs[tid] = s;
                rp[tid] = p;
                rq[tid] = q;
                rot[tid] = f;
            }
            // the angle threads are lanes 0..kJB-1 of wave 0: its ballot is the
            // round's rotation mask.  Slot r & 1 is rewritten two rounds
            // later, after a barrier every wave must reach once it has read
            // the slot.
            const unsigned long long bm = __ballot(f);
            if (tid == 0)
                rmask[r & 1] = (unsigned)bm;
            __syncthreads();
            // a round without rotations leaves S and U as they are
            if (rmask[r & 1] == 0u)
                continue;
            // the 2 x 2 block of pair k's rows and pair l's columns, for
            // thread (k, l): rows rotated by pair k, then columns by pair l
            // (a pair that does not rotate has c = 1, s = 0), and the
            // columns of U by pair l.  The blocks partition S and U.
            {
                const int k = tid >> 4, l = tid & 15;
                const unsigned m = rmask[r & 1];
                const bool rk = (m >> k) & 1u, rl = (m >> l) & 1u;
                if (rk || rl) {
                    const int pk = rp[k], qk = rq[k], pl = rp[l], ql = rq[l];
                    const double ck = rc[k], sk = rs[k], cl = rc[l], sl = rs[l];
                    const double a = S[pk * kJLD + pl], b = S[pk * kJLD + ql],
                                 c = S[qk * kJLD + pl], d = S[qk * kJLD + ql];
                    const double xa = ck * a - sk * c, xb = ck * b - sk * d;
                    const double xc = sk * a + ck * c, xd = sk * b + ck * d;
                    double na = cl * xa - sl * xb, nb = sl * xa + cl * xb;
                    double nc = cl * xc - sl * xd, nd = sl * xc + cl * xd;
                    if (k == l) // exact rotated 2 x 2 block of the pair
                        na = rpp[k], nb = 0.0, nc = 0.0, nd = rqq[k];
                    S[pk * kJLD + pl] = na;
                    S[pk * kJLD + ql] = nb;
                    S[qk * kJLD + pl] = nc;
                    S[qk * kJLD + ql] = nd;
                    if (rl) {
                        const double u0 = U[pk * kJLD + pl], u1 = U[pk * kJLD + ql],
                                     u2 = U[qk * kJLD + pl], u3 = U[qk * kJLD + ql];
                        U[pk * kJLD + pl] = cl * u0 - sl * u1;
                        U[pk * kJLD + ql] = sl * u0 + cl * u1;
                        U[qk * kJLD + pl] = cl * u2 - sl * u3;
                        U[qk * kJLD + ql] = sl * u2 + cl * u3;
                    }
                }
            }
            __syncthreads();
        }
        const int any = any_rot;
        __syncthreads(); // every thread has read any_rot before its reset
        if (!any)
            break;
        rotated = 1;
    }
    const uint64_t qo = bl.q_off + (uint64_t)tk.x * kJS * kJS;
    for (int e = tid; e < kJS * kJS; e += 256)
        Q[qo + e] = U[(e / kJS) * kJLD + (e % kJS)];
    if (tid == 0) {
        qflag[bl.f_off + tk.x] = rotated;
        if (rotated)
            atomicAdd(nrot, 1u);
    }
}

typedef double jv4d __attribute__((ext_vector_type(4)));

// C = X Y (TL: C = Y^T X) for kJS x kJS (32 x 32) operands held in LDS with
// row stride kJLD, on FP64 MFMA (16x16x4): wave w of the 256-thread
// workgroup computes the 16 x 16 quadrant of rows 16 (w >> 1) .., columns
// 16 (w & 1) ..; lane l feeds A[l % 16][l / 16] and B[l / 16][l % 16] per
// k-step, and accumulator element r of lane l is C[row][col] with
// (row, col) = mm_pos(r).
static_assert(kJS == 32, "the MFMA products assume 32 x 32 subproblems");
template <bool TL>
__device__ __forceinline__ jv4d lds_mm(const double *X, const double *Y) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int m0 = 16 * (w >> 1), n0 = 16 * (w & 1);
    jv4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < kJS / 4; ks++) {
        const int k = 4 * ks + lk;
        const double av = TL ? Y[k * kJLD + m0 + lr] : X[(m0 + lr) * kJLD + k];
        const double bv = TL ? X[k * kJLD + n0 + lr] : Y[k * kJLD + n0 + lr];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
    mfma_settle<jv4d, 1>(&acc);
    return acc;
}

// position of accumulator element r of this thread in the 32 x 32 result
__device__ __forceinline__ void mm_pos(int r, int &row, int &col) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    row = 16 * (w >> 1) + 4 * r + (lane >> 4);
    col = 16 * (w & 1) + (lane & 15);
}

// Region update A_PR <- Q_P^T A_PR Q_R (kind 1) and V slice update
// V[rows, P] <- V[rows, P] Q_P (kind 2) of one round.  Two LDS buffers:
// the operand and one Q; the product A_PR Q_R goes back into the operand
// buffer and the final result from registers to memory.
__global__ __launch_bounds__(256) void jacobi_update_kernel(
    const JacBlock *__restrict__ blocks, const JacTask *__restrict__ tasks,
    double *__restrict__ A, double *__restrict__ V,
    const double *__restrict__ Q, const int *__restrict__ qflag, int round) {
    __shared__ double X[kJS * kJLD];
    __shared__ double Y[kJS * kJLD];
    const JacTask tk = tasks[blockIdx.x];
    const JacBlock bl = blocks[tk.blk];
    const uint32_t N = bl.nb * kJB;
    const int r = round % (int)(bl.nb - 1);
    const int tid = threadIdx.x;
    if (tk.kind == 1) {
        const int fp = qflag[bl.f_off + tk.x], fr = qflag[bl.f_off + tk.y];
        if (!fp && !fr)
            return;
        int pi, pj, ri, rj;
        rr_pair((int)bl.nb, r, (int)tk.x, pi, pj);
        rr_pair((int)bl.nb, r, (int)tk.y, ri, rj);
        // every global operand into registers first (one memory latency):
        // element e = tid + 256 i of the region, of Q_R and of Q_P
        constexpr int EPT = kJS * kJS / 256;
        const uint64_t qro = bl.q_off + (uint64_t)tk.y * kJS * kJS;
        const uint64_t qpo = bl.q_off + (uint64_t)tk.x * kJS * kJS;
        double xa[EPT], qr[EPT], qp[EPT];
#pragma unroll
        for (int i = 0; i < EPT; i++) {
            const int e = tid + 256 * i, l = e / kJS, c = e % kJS;
            xa[i] = A[bl.a_off + (uint64_t)pair_index(pi, pj, l) * N +
                      pair_index(ri, rj, c)];
            qr[i] = fr ? Q[qro + e] : 0.0;
            qp[i] = fp ? Q[qpo + e] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < EPT; i++) {
            const int e = tid + 256 * i;
            X[(e / kJS) * kJLD + (e % kJS)] = xa[i];
        }
        jv4d acc;
        if (fr) {
#pragma unroll
            for (int i = 0; i < EPT; i++) {
                const int e = tid + 256 * i;
                Y[(e / kJS) * kJLD + (e % kJS)] = qr[i];
            }
            __syncthreads();
            acc = lds_mm<false>(X, Y); // X Q_R
            __syncthreads();
#pragma unroll
            for (int r4 = 0; r4 < 4; r4++) {
                int row, col;
                mm_pos(r4, row, col);
                X[row * kJLD + col] = acc[r4];
            }
        }
        if (fp) {
#pragma unroll
            for (int i = 0; i < EPT; i++) {
                const int e = tid + 256 * i;
                Y[(e / kJS) * kJLD + (e % kJS)] = qp[i];
            }
            __syncthreads();
            acc = lds_mm<true>(X, Y); // Q_P^T X
        } else {
            __syncthreads();
#pragma unroll
            for (int r4 = 0; r4 < 4; r4++) {
                int row, col;
                mm_pos(r4, row, col);
                acc[r4] = X[row * kJLD + col];
            }
        }
#pragma unroll
        for (int r4 = 0; r4 < 4; r4++) {
            int row, col;
            mm_pos(r4, row, col);
            A[bl.a_off + (uint64_t)pair_index(pi, pj, row) * N +
              pair_index(ri, rj, col)] = acc[r4];
        }
    } else {
        if (!qflag[bl.f_off + tk.x])
            return;
        int pi, pj;
        rr_pair((int)bl.nb, r, (int)tk.x, pi, pj);
        const uint32_t row0 = tk.y * kJS;
        constexpr int EPT = kJS * kJS / 256;
        const uint64_t qo = bl.q_off + (uint64_t)tk.x * kJS * kJS;
        double va[EPT], qp[EPT];
#pragma unroll
        for (int i = 0; i < EPT; i++) {
            const int e = tid + 256 * i, l = e / kJS, c = e % kJS;
            va[i] = V[bl.a_off + (uint64_t)(row0 + l) * N + pair_index(pi, pj, c)];
            qp[i] = Q[qo + e];
        }
#pragma unroll
        for (int i = 0; i < EPT; i++) {
            const int e = tid + 256 * i;
            X[(e / kJS) * kJLD + (e % kJS)] = va[i];
            Y[(e / kJS) * kJLD + (e % kJS)] = qp[i];
        }
        __syncthreads();
        const jv4d acc = lds_mm<false>(X, Y);
#pragma unroll
        for (int r4 = 0; r4 < 4; r4++) {
            int row, col;
            mm_pos(r4, row, col);
            V[bl.a_off + (uint64_t)(row0 + row) * N + pair_index(pi, pj, col)] =
                acc[r4];
        }
    }
}

// Eigenvalues as Rayleigh quotients against the original matrix:
// w[p] = sum_i V[i][p] Y[i][p] / sum_i V[i][p]^2, Y = A_orig V (compact).
// The diagonal left by the block updates carries an absolute error of a
// few eps ||A|| per update; the quotient of the converged vector is
// accurate to O(eps ||A||) with the vector's own error entering
// quadratically.  Workgroup (g, b): columns 64 g .. 64 g + 63 of matrix b,
// four row slices per column summed in LDS in a fixed order.
__global__ __launch_bounds__(256) void rayleigh_kernel(
    const JacBlock *__restrict__ blocks, const uint64_t *__restrict__ w_off,
    const uint64_t *__restrict__ y_off, const double *__restrict__ V,
    const double *__restrict__ Y, double *__restrict__ w) {
    __shared__ double sn[4][64], sd[4][64];
    const int b = blockIdx.y;
    const JacBlock bl = blocks[b];
    const uint32_t N = bl.nb * kJB;
    const uint32_t p = blockIdx.x * 64 + (threadIdx.x & 63);
    const int sl = threadIdx.x >> 6;
    double s = 0.0, nn = 0.0;
    if (p < bl.n)
        for (uint32_t i = sl; i < bl.n; i += 4) {
            const double v = V[bl.a_off + (uint64_t)i * N + p];
            s = fma(v, Y[y_off[b] + (uint64_t)i * bl.n + p], s);
            nn = fma(v, v, nn);
        }
    sn[sl][threadIdx.x & 63] = s;
    sd[sl][threadIdx.x & 63] = nn;
    __syncthreads();
    if (sl == 0 && p < bl.n) {
        const int c = threadIdx.x;
        w[w_off[b] + p] = ((sn[0][c] + sn[1][c]) + (sn[2][c] + sn[3][c])) /
                          ((sd[0][c] + sd[1][c]) + (sd[2][c] + sd[3][c]));
    }
}

// One workgroup per eigenvector: out row j of matrix b = V column perm[j],
// scaled to unit 2-norm (the accumulated rotations leave norms off by
// O(sweeps * eps)) with the sign that makes its largest-magnitude
// component positive (the first such component on ties).
struct VecJob {
    uint32_t blk, col;
    uint64_t dst; // row start in the output
};
__global__ __launch_bounds__(256) void eigvec_rows_kernel(
    const JacBlock *__restrict__ blocks, const VecJob *__restrict__ jobs,
    const double *__restrict__ V, double *__restrict__ out) {
    __shared__ double sv[256], sq[256];
    __shared__ int si[256];
    const VecJob jb = jobs[blockIdx.x];
    const JacBlock bl = blocks[jb.blk];
    const uint32_t N = bl.nb * kJB;
    double best = -1.0, ss = 0.0;
    int bi = 0x7FFFFFFF;
    for (uint32_t i = threadIdx.x; i < bl.n; i += 256) {
        const double v = V[bl.a_off + (uint64_t)i * N + jb.col];
        ss = fma(v, v, ss);
        if (fabs(v) > best)
            best = fabs(v), bi = (int)i;
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    sq[threadIdx.x] = ss;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double o = sv[threadIdx.x + w];
            const int oi = si[threadIdx.x + w];
            if (o > sv[threadIdx.x] ||
                (o == sv[threadIdx.x] && oi < si[threadIdx.x]))
                sv[threadIdx.x] = o, si[threadIdx.x] = oi;
            sq[threadIdx.x] += sq[threadIdx.x + w];
        }
        __syncthreads();
    }
    const double sgn =
        (V[bl.a_off + (uint64_t)si[0] * N + jb.col] < 0.0 ? -1.0 : 1.0) /
        sqrt(sq[0]);
    for (uint32_t i = threadIdx.x; i < bl.n; i += 256)
        out[jb.dst + i] = sgn * V[bl.a_off + (uint64_t)i * N + jb.col];
}

// copies of eigenvector rows into the rotation matrix: row `src` (length
// n) to dst + i * stride for i < n
struct RowCopy {
    uint64_t src, dst;
    uint32_t n, stride;
};
__global__ __launch_bounds__(256) void row_copy_kernel(
    const RowCopy *__restrict__ jobs, int njobs, const double *__restrict__ E,
    double *__restrict__ out) {
    for (int j = blockIdx.x; j < njobs; j += gridDim.x) {
        const RowCopy c = jobs[j];
        for (uint32_t i = threadIdx.x; i < c.n; i += 256)
            out[c.dst + (uint64_t)i * c.stride] = E[c.src + i];
    }
}

} // namespace

// ----------------------------------------------------------------- solver
namespace {

// grouped-GEMM tasks of one truncation product, uploaded in one allocation
// Density-matrix noise: D_b[i][j] += coef[b] T_b[i][j] on the lower
// triangle (i >= j), mirrored to the upper one, so that the matrix handed
// to the eigensolver is the symmetric matrix of its lower triangle.  Upper
// entries are only written, never read: no thread depends on another.
__global__ __launch_bounds__(256) void dm_noise_kernel(
    const uint64_t *__restrict__ off, const uint32_t *__restrict__ dim,
    const double *__restrict__ coef, int nb, const double *__restrict__ T,
    double *__restrict__ D) {
    for (int b = blockIdx.y; b < nb; b += gridDim.y) {
        const uint32_t n = dim[b];
        const uint64_t o = off[b], nn = (uint64_t)n * n;
        const double c = coef[b];
        for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < nn;
             e += (uint64_t)gridDim.x * 256) {
            const uint32_t i = (uint32_t)(e / n), j = (uint32_t)(e % n);
            if (j > i)
                continue;
            const double v = D[o + e] + c * T[o + e];
            D[o + e] = v;
            D[o + (uint64_t)j * n + i] = v;
        }
    }
}

struct TaskSet {
    std::vector<Task> t[kTileConfigs];
    std::vector<Seg> s;
    void *dev = nullptr;
    Task *d_t[kTileConfigs] = {};
    Seg *d_s = nullptr;
    ~TaskSet() {
        if (dev)
            hipFree(dev);
    }
    int upload(b2g_ctx ctx) {
        for (int c = 0; c < kTileConfigs; c++)
            sort_tasks(t[c], s);
        size_t bytes = s.size() * sizeof(Seg);
        for (int c = 0; c < kTileConfigs; c++)
            bytes += t[c].size() * sizeof(Task);
        char *host = nullptr;
        int r = ctx_stage(ctx, bytes + 16, &host);
        if (r != B2G_OK)
            return r;
        B2G_HIP(hipMalloc(&dev, bytes + 16));
        size_t off = 0;
        auto put = [&](const void *src, size_t nb) {
            if (nb)
                memcpy(host + off, src, nb);
            char *d = (char *)dev + off;
            off += nb;
            return d;
        };
        d_s = (Seg *)put(s.data(), s.size() * sizeof(Seg));
        for (int c = 0; c < kTileConfigs; c++)
            d_t[c] = (Task *)put(t[c].data(), t[c].size() * sizeof(Task));
        B2G_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice,
                               ctx->stream));
        B2G_HIP(hipStreamSynchronize(ctx->stream));
        return B2G_OK;
    }
    int run(b2g_ctx ctx, const double *A, const double *B, double *D) const {
        for (int c = 0; c < kTileConfigs; c++)
            if (!t[c].empty())
                B2G_HIP((launch_seg_gemm<double, double, double, double, 5>(
                    c, A, B, D, D, d_t[c], (int)t[c].size(), d_s, 0, 0, 0, 0,
                    1.0, 0, 1, ctx->num_cu * 8, ctx->stream)));
        return B2G_OK;
    }
};

} // namespace

// Batched symmetric eigensolver state (device buffers of one call).
struct Syev {
    std::vector<JacBlock> blk;
    std::vector<uint64_t> w_off; // compact eigenvalue offsets
    uint64_t a_len = 0, q_len = 0, f_len = 0, w_len = 0;
    uint32_t max_nb = 2;
    double *A = nullptr, *V = nullptr, *Q = nullptr, *w = nullptr;
    int *qflag = nullptr;
    unsigned *nrot = nullptr, *h_nrot = nullptr; // device counter, pinned copy
    JacBlock *d_blk = nullptr;
    JacTask *d_pair = nullptr, *d_upd = nullptr;
    uint64_t *d_src_off = nullptr, *d_w_off = nullptr;
    size_t n_pair = 0, n_upd = 0;
    int sweeps = 0;
    ~Syev() {
        for (void *p : {(void *)A, (void *)V, (void *)Q, (void *)w,
                        (void *)qflag, (void *)nrot, (void *)d_blk,
                        (void *)d_pair, (void *)d_upd, (void *)d_src_off,
                        (void *)d_w_off})
            if (p)
                hipFree(p);
        if (h_nrot)
            hipHostFree(h_nrot);
    }
};

// Eigen-decomposition of the compact symmetric matrices src + src_off[i]
// (order n[i]): eigenvalues (ascending) into w_sorted (host, compact),
// eigenvector rows (same order, signed) into E (device, src's layout).
static int syev_run(b2g_ctx ctx, const std::vector<uint32_t> &n,
                    const std::vector<uint64_t> &src_off, const double *src,
                    double *E, std::vector<double> &w_sorted, int *sweeps) {
    hipStream_t st = ctx->stream;
    Syev s;
    const size_t nm = n.size();
    s.w_off.resize(nm + 1, 0);
    for (size_t i = 0; i < nm; i++) {
        JacBlock b;
        b.nb = std::max<uint32_t>(2, (n[i] + kJB - 1) / kJB);
        b.nb += b.nb & 1;
        const uint64_t N = (uint64_t)b.nb * kJB;
        // the Rayleigh-quotient product addresses the original block and
        // the padded eigenvector block with 32-bit byte offsets from each
        // block's start (2^31 bytes): both must hold fewer than 2^28
        // elements
        B2G_REQUIRE(N * N < ((uint64_t)1 << 28), B2G_ERR_ARG,
                    "eigensolver: matrix order above " +
                        std::to_string(kMaxSyevOrder));
        b.n = n[i];
        b.a_off = s.a_len;
        b.q_off = s.q_len;
        b.f_off = (uint32_t)s.f_len;
        b.pad = 0;
        b.floor_abs = 0.0;
        s.a_len += N * N;
        s.q_len += (uint64_t)(b.nb / 2) * kJS * kJS;
        s.f_len += b.nb / 2;
        s.max_nb = std::max(s.max_nb, b.nb);
        s.w_off[i + 1] = s.w_off[i] + n[i];
        s.blk.push_back(b);
    }
    s.w_len = s.w_off[nm];
    w_sorted.assign(s.w_len, 0.0);
    if (nm == 0)
        return B2G_OK;
    B2G_REQUIRE(nm <= 65535, B2G_ERR_ARG,
                "eigensolver: more than 65535 matrices");
    // task lists: pairs, then regions and V slices
    std::vector<JacTask> pairs, upd;
    for (size_t i = 0; i < nm; i++) {
        const uint32_t h = s.blk[i].nb / 2;
        for (uint32_t k = 0; k < h; k++)
            pairs.push_back(JacTask{(uint32_t)i, 0, k, 0});
        for (uint32_t a = 0; a < h; a++)
            for (uint32_t b = 0; b < h; b++)
                upd.push_back(JacTask{(uint32_t)i, 1, a, b});
        for (uint32_t a = 0; a < h; a++)
            for (uint32_t rb = 0; rb < h; rb++) // slices of kJS rows
                upd.push_back(JacTask{(uint32_t)i, 2, a, rb});
    }
    s.n_pair = pairs.size();
    s.n_upd = upd.size();
    B2G_REQUIRE(s.n_pair < (1u << 31) && s.n_upd < (1u << 31), B2G_ERR_ARG,
                "eigensolver: too many blocks");
    B2G_HIP(hipMalloc(&s.A, s.a_len * sizeof(double)));
    B2G_HIP(hipMalloc(&s.V, s.a_len * sizeof(double)));
    B2G_HIP(hipMalloc(&s.Q, s.q_len * sizeof(double)));
    B2G_HIP(hipMalloc(&s.w, std::max<uint64_t>(s.w_len, 1) * sizeof(double)));
    B2G_HIP(hipMalloc(&s.qflag, s.f_len * sizeof(int)));
    B2G_HIP(hipMalloc(&s.d_blk, nm * sizeof(JacBlock)));
    B2G_HIP(hipMalloc(&s.d_pair, s.n_pair * sizeof(JacTask)));
    B2G_HIP(hipMalloc(&s.d_upd, s.n_upd * sizeof(JacTask)));
    B2G_HIP(hipMalloc(&s.d_src_off, nm * sizeof(uint64_t)));
    B2G_HIP(hipMalloc(&s.d_w_off, nm * sizeof(uint64_t)));
    B2G_HIP(hipMalloc(&s.nrot, sizeof(unsigned)));
    B2G_HIP(hipHostMalloc((void **)&s.h_nrot, sizeof(unsigned)));
    B2G_HIP(hipMemcpyAsync(s.d_blk, s.blk.data(), nm * sizeof(JacBlock),
                           hipMemcpyHostToDevice, st));
    B2G_HIP(hipMemcpyAsync(s.d_pair, pairs.data(), s.n_pair * sizeof(JacTask),
                           hipMemcpyHostToDevice, st));
    B2G_HIP(hipMemcpyAsync(s.d_upd, upd.data(), s.n_upd * sizeof(JacTask),
                           hipMemcpyHostToDevice, st));
    B2G_HIP(hipMemcpyAsync(s.d_src_off, src_off.data(), nm * sizeof(uint64_t),
                           hipMemcpyHostToDevice, st));
    B2G_HIP(hipMemcpyAsync(s.d_w_off, s.w_off.data(), nm * sizeof(uint64_t),
                           hipMemcpyHostToDevice, st));
    {
        const uint64_t nn = (uint64_t)s.max_nb * kJB * s.max_nb * kJB;
        const unsigned gx = (unsigned)std::min<uint64_t>((nn + 255) / 256, 64);
        const unsigned gy = (unsigned)std::min<size_t>(nm, 4096);
        hipLaunchKernelGGL(jacobi_init_kernel, dim3(gx, gy), dim3(256), 0, st,
                           s.d_blk, (int)nm, s.d_src_off, src, s.A, s.V);
        B2G_HIP(hipGetLastError());
        const unsigned nparts = 32;
        double *part = nullptr;
        B2G_HIP(hipMalloc(&part, nm * nparts * sizeof(double)));
        std::unique_ptr<double, int (*)(void *)> part_guard(
            part, [](void *p) { return (int)hipFree(p); });
        hipLaunchKernelGGL(jacobi_norm_kernel, dim3(nparts, (unsigned)nm),
                           dim3(256), 0, st, s.d_blk, s.A, part);
        B2G_HIP(hipGetLastError());
        hipLaunchKernelGGL(jacobi_floor_kernel,
                           dim3((unsigned)((nm + 255) / 256)), dim3(256), 0,
                           st, s.d_blk, (int)nm, part, (int)nparts);
        B2G_HIP(hipGetLastError());
        B2G_HIP(hipStreamSynchronize(st));
    }
    const int max_sweeps = 60;
    const double tol = 2.220446049250313e-16;
    // cyclic sweeps per pair subproblem and round (B2G_JACOBI_INNER): the
    // outer iteration converges with partially diagonalised subproblems,
    // at a fraction of the LDS work per round
    int inner = 1;
    if (const char *e = getenv("B2G_JACOBI_INNER"))
        if (atoi(e) > 0)
            inner = atoi(e);
    int sweep = 0;
    for (; sweep < max_sweeps; sweep++) {
        B2G_HIP(hipMemsetAsync(s.nrot, 0, sizeof(unsigned), st));
        for (uint32_t r = 0; r + 1 < s.max_nb; r++) {
            hipLaunchKernelGGL(jacobi_pair_kernel, dim3((unsigned)s.n_pair),
                               dim3(256), 0, st, s.d_blk, s.d_pair, s.A, s.Q,
                               s.qflag, (int)r, tol, inner, s.nrot);
            B2G_HIP(hipGetLastError());
            hipLaunchKernelGGL(jacobi_update_kernel, dim3((unsigned)s.n_upd),
                               dim3(256), 0, st, s.d_blk, s.d_upd, s.A,
                               s.V, s.Q, s.qflag, (int)r);
            B2G_HIP(hipGetLastError());
        }
        B2G_HIP(hipMemcpyAsync(s.h_nrot, s.nrot, sizeof(unsigned),
                               hipMemcpyDeviceToHost, st));
        B2G_HIP(hipStreamSynchronize(st));
        if (getenv("B2G_JACOBI_TRACE") != nullptr)
            fprintf(stderr, "b2g jacobi sweep %d: %u rotating pair subproblems\n",
                    sweep, *s.h_nrot);
        if (*s.h_nrot == 0)
            break;
    }
    if (sweeps)
        *sweeps = sweep + 1;
    B2G_REQUIRE(sweep < max_sweeps, B2G_ERR_NOCONV,
                "eigensolver: block Jacobi did not converge");
    // Y = A_orig V[:, :n] per matrix on the grouped GEMM, then the
    // quotients
    {
        TaskSet ts;
        for (size_t i = 0; i < nm; i++) {
            if (n[i] == 0)
                continue;
            Seg sg;
            sg.a_off = src_off[i], sg.lda = n[i];
            sg.b_off = s.blk[i].a_off, sg.ldb = s.blk[i].nb * kJB;
            sg.K = n[i], sg.flags = 0, sg.alpha = 1.0;
            const uint32_t sb = (uint32_t)ts.s.size();
            ts.s.push_back(sg);
            add_tiles(ts.t, src_off[i], n[i], n[i], n[i], sb, sb + 1, 0);
        }
        uint64_t y_len = 0;
        for (size_t i = 0; i < nm; i++)
            y_len = std::max<uint64_t>(y_len, src_off[i] + (uint64_t)n[i] * n[i]);
        double *Y = nullptr;
        B2G_HIP(hipMalloc(&Y, std::max<uint64_t>(y_len, 1) * sizeof(double)));
        std::unique_ptr<double, int (*)(void *)> y_guard(
            Y, [](void *p) { return (int)hipFree(p); });
        int r = ts.upload(ctx);
        if (r == B2G_OK)
            r = ts.run(ctx, src, s.V, Y);
        if (r != B2G_OK)
            return r;
        unsigned gx = 1;
        for (size_t i = 0; i < nm; i++)
            gx = std::max<unsigned>(gx, (n[i] + 63) / 64);
        hipLaunchKernelGGL(rayleigh_kernel, dim3(gx, (unsigned)nm), dim3(256),
                           0, st, s.d_blk, s.d_w_off, s.d_src_off, s.V, Y,
                           s.w);
        B2G_HIP(hipGetLastError());
        B2G_HIP(hipStreamSynchronize(st));
    }
    std::vector<double> w(s.w_len);
    B2G_HIP(hipMemcpyAsync(w.data(), s.w, s.w_len * sizeof(double),
                           hipMemcpyDeviceToHost, st));
    B2G_HIP(hipStreamSynchronize(st));
    // ascending order per matrix (positions in order among equal values)
    std::vector<VecJob> jobs;
    jobs.reserve(s.w_len);
    for (size_t i = 0; i < nm; i++) {
        std::vector<std::pair<double, uint32_t>> o(n[i]);
        for (uint32_t j = 0; j < n[i]; j++)
            o[j] = std::make_pair(w[s.w_off[i] + j], j);
        std::sort(o.begin(), o.end());
        for (uint32_t j = 0; j < n[i]; j++) {
            w_sorted[s.w_off[i] + j] = o[j].first;
            jobs.push_back(VecJob{(uint32_t)i, o[j].second,
                                  src_off[i] + (uint64_t)j * n[i]});
        }
    }
    if (!jobs.empty()) {
        VecJob *d_jobs = nullptr;
        B2G_HIP(hipMalloc(&d_jobs, jobs.size() * sizeof(VecJob)));
        hipError_t e = hipMemcpyAsync(d_jobs, jobs.data(),
                                      jobs.size() * sizeof(VecJob),
                                      hipMemcpyHostToDevice, st);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(eigvec_rows_kernel, dim3((unsigned)jobs.size()),
                               dim3(256), 0, st, s.d_blk, d_jobs, s.V, E);
            e = hipGetLastError();
        }
        const hipError_t e2 = hipStreamSynchronize(st);
        hipFree(d_jobs);
        B2G_HIP(e);
        B2G_HIP(e2);
    }
    return B2G_OK;
}

} // namespace b2g

using namespace b2g;

extern "C" int b2g_syev_batched(b2g_ctx ctx, int n_mats, const uint32_t *n,
                                const uint64_t *off, const double *a,
                                double *w, double *v, int *sweeps) {
    B2G_REQUIRE(ctx && (n_mats == 0 || (n && off && a && w && v)), B2G_ERR_ARG,
                "NULL arg");
    B2G_REQUIRE(n_mats >= 0, B2G_ERR_ARG, "negative matrix count");
    B2G_HIP(hipSetDevice(ctx->device));
    std::vector<uint32_t> nv(n, n + n_mats);
    std::vector<uint64_t> ov(off, off + n_mats);
    for (int i = 0; i < n_mats; i++)
        B2G_REQUIRE(nv[i] <= (uint32_t)kMaxSyevOrder, B2G_ERR_ARG,
                    "matrix order above " + std::to_string(kMaxSyevOrder));
    std::vector<double> ws;
    const int r = syev_run(ctx, nv, ov, a, v, ws, sweeps);
    if (r != B2G_OK)
        return r;
    if (!ws.empty())
        B2G_HIP(hipMemcpyAsync(w, ws.data(), ws.size() * sizeof(double),
                               hipMemcpyHostToDevice, ctx->stream));
    B2G_HIP(hipStreamSynchronize(ctx->stream));
    return B2G_OK;
}

// ------------------------------------------------------------------ split
// block2's generator (Random::rng, src/core/utils.hpp:196-252) as a
// handle whose stream continues across calls
struct b2g_rng_s {
    std::mt19937 gen;
};

struct b2g_split_s {
    b2g_ctx ctx = nullptr;
    int n_roots = 1; // the wavefunction part holds n_roots consecutive roots
    double error = 0.0;
    std::vector<double> eigvals; // per density-matrix block, ascending
    std::vector<double> spectra; // sqrt(max(0, w)) (store_wfn_spectra)
    struct Part {
        uint64_t dq = 0;
        int fermion = 0;
        std::vector<uint64_t> quanta;
        std::vector<uint32_t> nb, nk;
        std::vector<uint64_t> shift;
        double *data = nullptr;
        uint64_t len = 0;
    };
    Part part[3]; // rotation-side left, right, density matrix
    ~b2g_split_s() {
        for (Part &p : part)
            if (p.data)
                hipFree(p.data);
    }
};

namespace b2g {

namespace {

int find_q(const std::vector<uint64_t> &qs, uint64_t q) {
    auto it = std::lower_bound(qs.begin(), qs.end(), q);
    return it != qs.end() && *it == q ? (int)(it - qs.begin()) : -1;
}

// sort_blocks (src/core/sparse_matrix.hpp:790-): blocks by quanta
void sort_part(b2g_split_s::Part &p) {
    std::vector<size_t> idx(p.quanta.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&p](size_t a, size_t b) {
        return p.quanta[a] < p.quanta[b];
    });
    b2g_split_s::Part q;
    for (size_t i : idx) {
        q.quanta.push_back(p.quanta[i]);
        q.nb.push_back(p.nb[i]);
        q.nk.push_back(p.nk[i]);
    }
    p.quanta.swap(q.quanta);
    p.nb.swap(q.nb);
    p.nk.swap(q.nk);
}

void set_shifts(b2g_split_s::Part &p) {
    p.shift.assign(p.quanta.size(), 0);
    uint64_t off = 0;
    for (size_t i = 0; i < p.quanta.size(); i++) {
        p.shift[i] = off;
        off += (uint64_t)p.nb[i] * p.nk[i];
    }
    p.len = off;
}

} // namespace

} // namespace b2g

namespace b2g {

// OperatorFunctions::trans_product's noise (src/core/operator_functions.hpp:
// 776-835, called by MovingEnvironment::density_matrix with sqrt(noise),
// moving_environment.hpp:3554-3577), with block2's generator
// (Random::fill, src/core/utils.hpp:226-252): std::mt19937 seeded with
// `seed`, one uniform_real_distribution<double> draw per element in storage
// order.
//   wavefunction:   T over the wavefunction's layout, uniform in
//                   [-0.5, 0.5); D += (noise / |T|^2) T T^T per wavefunction
//                   block (the density-matrix products of T)
//   density matrix: T over the density matrix's layout, uniform in [0, 1);
//                   D_b += (noise / |T|) T_b once per wavefunction block of
//                   sector b; the eigensolver then sees the lower triangle
//                   (block2's xsyev "U" on row-major storage)
// State-averaged roots (density_matrix_with_multi_target,
// moving_environment.hpp:3582-3608) each draw their own T from the
// continuing stream, in root order.
static int add_noise(b2g_ctx ctx, const b2g_sparse_info *wfn, uint64_t psi_len,
                     int n_roots, const b2g_split_s::Part &dm,
                     const std::vector<int> &wfn_dm, const TaskSet &ts,
                     double noise, int noise_type, std::mt19937 &gen) {
    (void)wfn;
    const bool wav = noise_type == B2G_NOISE_WAVEFUNCTION;
    // one draw per root from the continuing stream (every root's
    // trans_product draws its own tensor)
    const uint64_t len = wav ? psi_len : dm.len;
    if (len == 0)
        return B2G_OK;
    std::vector<double> T(len * n_roots), scal(n_roots);
    {
        std::uniform_real_distribution<double> distr(wav ? -0.5 : 0.0,
                                                     wav ? 0.5 : 1.0);
        for (uint64_t i = 0; i < len * n_roots; i++)
            T[i] = distr(gen);
    }
    for (int r = 0; r < n_roots; r++) {
        double ss = 0.0;
        for (uint64_t i = 0; i < len; i++)
            ss += T[r * len + i] * T[r * len + i];
        B2G_REQUIRE(ss > 0.0, B2G_ERR_ARG, "noise: zero random tensor");
        scal[r] = wav ? noise / ss : noise / std::sqrt(ss);
    }
    if (!wav) {
        // the roots' noise summed on the host: sum_r scal[r] T_r
        for (uint64_t i = 0; i < len; i++) {
            double v = scal[0] * T[i];
            for (int r = 1; r < n_roots; r++)
                v += scal[r] * T[r * len + i];
            T[i] = v;
        }
        T.resize(len);
    }
    double *dT = nullptr;
    B2G_HIP(hipMalloc(&dT, T.size() * sizeof(double)));
    std::unique_ptr<double, int (*)(void *)> t_guard(
        dT, [](void *p) { return (int)hipFree(p); });
    B2G_HIP(hipMemcpy(dT, T.data(), T.size() * sizeof(double),
                      hipMemcpyHostToDevice));
    if (wav) {
        // the same products as the density matrix, on each root's T
        // (segments of root r start at r * psi_len), accumulated
        TaskSet tn;
        tn.s = ts.s;
        for (Seg &sg : tn.s)
            sg.alpha = scal[(int)(sg.a_off / psi_len)];
        for (int c = 0; c < kTileConfigs; c++) {
            tn.t[c] = ts.t[c];
            for (Task &t : tn.t[c])
                t.beta |= 1u;
        }
        int r = tn.upload(ctx);
        if (r == B2G_OK)
            r = tn.run(ctx, dT, dT, dm.data);
        if (r != B2G_OK)
            return r;
        B2G_HIP(hipStreamSynchronize(ctx->stream));
        return B2G_OK;
    }
    const int nd = (int)dm.quanta.size();
    std::vector<double> coef(nd, 0.0);
    for (size_t i = 0; i < wfn_dm.size(); i++)
        coef[wfn_dm[i]] += 1.0;
    std::vector<uint32_t> dim(dm.nb.begin(), dm.nb.end());
    const size_t bytes = nd * (sizeof(uint64_t) + sizeof(uint32_t) + sizeof(double));
    char *dbuf = nullptr;
    B2G_HIP(hipMalloc(&dbuf, bytes));
    std::unique_ptr<char, int (*)(void *)> b_guard(
        dbuf, [](void *p) { return (int)hipFree(p); });
    uint64_t *d_off = (uint64_t *)dbuf;
    double *d_coef = (double *)(dbuf + nd * sizeof(uint64_t));
    uint32_t *d_dim = (uint32_t *)(dbuf + nd * (sizeof(uint64_t) + sizeof(double)));
    B2G_HIP(hipMemcpy(d_off, dm.shift.data(), nd * sizeof(uint64_t),
                      hipMemcpyHostToDevice));
    B2G_HIP(hipMemcpy(d_coef, coef.data(), nd * sizeof(double),
                      hipMemcpyHostToDevice));
    B2G_HIP(hipMemcpy(d_dim, dim.data(), nd * sizeof(uint32_t),
                      hipMemcpyHostToDevice));
    hipLaunchKernelGGL(dm_noise_kernel, dim3(16, (unsigned)std::min(nd, 4096)),
                       dim3(256), 0, ctx->stream, d_off, d_dim, d_coef, nd, dT,
                       dm.data);
    B2G_HIP(hipGetLastError());
    B2G_HIP(hipStreamSynchronize(ctx->stream));
    return B2G_OK;
}

// Extra density-matrix terms: vectors of their own block layouts, in the
// same device buffer as the wavefunction (at `offs`), each adding
// v v^T (trace_right) or v^T v to the density-matrix sectors of its blocks.
struct ExtraTerms {
    int n = 0;
    const b2g_sparse_info *infos = nullptr;
    const uint64_t *offs = nullptr;
};

static int split_impl(b2g_ctx ctx, int symmetry, const b2g_sparse_info *wfn,
                      const double *psi, int n_roots, const double *weights,
                      int k, double cutoff, int trace_right, int normalize,
                      unsigned trunc_type, double scale, double noise,
                      int noise_type, std::mt19937 *gen, b2g_split *out,
                      const ExtraTerms &extra = ExtraTerms()) {
    B2G_REQUIRE(ctx && wfn && out && (psi || wfn->n == 0), B2G_ERR_ARG,
                "NULL arg");
    B2G_REQUIRE(n_roots >= 1 && weights != nullptr, B2G_ERR_ARG,
                "need at least one root and its weight");
    B2G_REQUIRE(symmetry == B2G_SZ || symmetry == B2G_SU2, B2G_ERR_ARG,
                "unknown symmetry");
    B2G_REQUIRE(wfn->n >= 0 && (wfn->n == 0 || (wfn->quanta &&
                                                wfn->n_states_bra &&
                                                wfn->n_states_ket &&
                                                wfn->block_shifts)),
                B2G_ERR_ARG, "bad wavefunction info");
    B2G_REQUIRE(k >= -1, B2G_ERR_ARG, "k must be -1 (keep all) or >= 0");
    B2G_REQUIRE(noise == 0.0 || ((noise_type == B2G_NOISE_WAVEFUNCTION ||
                                  noise_type == B2G_NOISE_DENSITY_MATRIX) &&
                                 noise > 0.0),
                B2G_ERR_ARG,
                "noise: positive, of type wavefunction or density matrix");
    *out = nullptr;
    B2G_HIP(hipSetDevice(ctx->device));
    SetupTimer tm("split");
    tile_env_refresh();
    const Symm sy = symmetry == B2G_SZ ? Symm::SZ : Symm::SU2;
    const int nw = wfn->n;
    const uint64_t wdq = wfn->delta_quantum;
    std::unique_ptr<b2g_split_s> h(new b2g_split_s());
    h->ctx = ctx;
    h->n_roots = n_roots;
    // ---- density-matrix info (SparseMatrixInfo::initialize_dm)
    std::vector<uint64_t> dq_of(nw);
    for (int i = 0; i < nw; i++)
        dq_of[i] = trace_right ? q_get_bra(sy, wfn->quanta[i], wdq)
                               : q_neg(sy, q_get_ket(sy, wfn->quanta[i]));
    b2g_split_s::Part &dm = h->part[2];
    dm.quanta = dq_of;
    std::sort(dm.quanta.begin(), dm.quanta.end());
    dm.quanta.erase(std::unique(dm.quanta.begin(), dm.quanta.end()),
                    dm.quanta.end());
    const int nd = (int)dm.quanta.size();
    dm.nb.assign(nd, 0);
    dm.nk.assign(nd, 0);
    dm.dq = 0; // vacuum
    std::vector<int> wfn_dm(nw);
    for (int i = 0; i < nw; i++) {
        wfn_dm[i] = find_q(dm.quanta, dq_of[i]);
        const uint32_t d = trace_right ? wfn->n_states_bra[i]
                                       : wfn->n_states_ket[i];
        dm.nb[wfn_dm[i]] = dm.nk[wfn_dm[i]] = d;
        const uint64_t sz = (uint64_t)wfn->n_states_bra[i] * wfn->n_states_ket[i];
        B2G_REQUIRE(sz < ((uint64_t)1 << 28), B2G_ERR_ARG,
                    "wavefunction block of 2^28 elements or more");
    }
    // elements of one root (roots are stored one after the other)
    uint64_t psi_len = 0;
    for (int i = 0; i < nw; i++)
        psi_len = std::max<uint64_t>(
            psi_len, wfn->block_shifts[i] + (uint64_t)wfn->n_states_bra[i] *
                                                wfn->n_states_ket[i]);
    for (int i = 0; i < nw; i++) {
        const uint32_t d = trace_right ? wfn->n_states_bra[i]
                                       : wfn->n_states_ket[i];
        B2G_REQUIRE(dm.nb[wfn_dm[i]] == d, B2G_ERR_ARG,
                    "wavefunction blocks of one density-matrix sector differ "
                    "in dimension");
    }
    set_shifts(dm);
    B2G_REQUIRE(dm.len < ((uint64_t)1 << 40), B2G_ERR_ARG,
                "density matrix too large");
    // ---- density matrix on the device: per block, one task list over the
    // wavefunction blocks of its sector (trans_product, scale = factor^2)
    {
        TaskSet ts;
        std::vector<std::vector<int>> by(nd);
        for (int i = 0; i < nw; i++)
            by[wfn_dm[i]].push_back(i);
        // extra terms' blocks by density-matrix sector: (term, block)
        std::vector<std::vector<std::pair<int, int>>> xby(nd);
        for (int e = 0; e < extra.n; e++) {
            const b2g_sparse_info &xi = extra.infos[e];
            B2G_REQUIRE(xi.n >= 0 && (xi.n == 0 || (xi.quanta &&
                                                    xi.n_states_bra &&
                                                    xi.n_states_ket &&
                                                    xi.block_shifts)),
                        B2G_ERR_ARG, "bad extra term info");
            for (int j = 0; j < xi.n; j++) {
                const uint64_t q =
                    trace_right
                        ? q_get_bra(sy, xi.quanta[j], xi.delta_quantum)
                        : q_neg(sy, q_get_ket(sy, xi.quanta[j]));
                const int b = find_q(dm.quanta, q);
                if (b < 0) // TraceProduct skips sectors the DM lacks
                    continue;
                const uint32_t d = trace_right ? xi.n_states_bra[j]
                                               : xi.n_states_ket[j];
                B2G_REQUIRE(d == dm.nb[b], B2G_ERR_ARG,
                            "extra term block of another sector dimension");
                xby[b].emplace_back(e, j);
            }
        }
        for (int b = 0; b < nd; b++) {
            const uint32_t sb = (uint32_t)ts.s.size();
            uint64_t K = 0;
            for (int rt = 0; rt < n_roots; rt++)
            for (int i : by[b]) {
                const uint32_t nbr = wfn->n_states_bra[i],
                               nkt = wfn->n_states_ket[i];
                Seg s;
                s.a_off = (uint64_t)rt * psi_len + wfn->block_shifts[i];
                s.b_off = s.a_off;
                // DensityMatrix of root rt with factor sqrt(w scale)
                s.alpha = weights[rt] * scale;
                if (trace_right) { // psi psi^T: K = ket dimension
                    s.lda = nkt, s.ldb = nkt, s.K = nkt, s.flags = 2;
                } else { // psi^T psi: K = bra dimension
                    s.lda = nkt, s.ldb = nkt, s.K = nbr, s.flags = 1;
                }
                if (s.K == 0)
                    continue;
                K += s.K;
                ts.s.push_back(s);
            }
            for (const auto &ej : xby[b]) {
                const b2g_sparse_info &xi = extra.infos[ej.first];
                const uint32_t nbr = xi.n_states_bra[ej.second],
                               nkt = xi.n_states_ket[ej.second];
                Seg s;
                s.a_off = extra.offs[ej.first] + xi.block_shifts[ej.second];
                s.b_off = s.a_off;
                s.alpha = 1.0;
                if (trace_right)
                    s.lda = nkt, s.ldb = nkt, s.K = nkt, s.flags = 2;
                else
                    s.lda = nkt, s.ldb = nkt, s.K = nbr, s.flags = 1;
                if (s.K == 0)
                    continue;
                K += s.K;
                ts.s.push_back(s);
            }
            if (dm.nb[b] > 0)
                add_tiles(ts.t, dm.shift[b], dm.nb[b], dm.nb[b], K, sb,
                          (uint32_t)ts.s.size(), 0);
        }
        B2G_HIP(hipMalloc(&dm.data, std::max<uint64_t>(dm.len, 1) *
                                        sizeof(double)));
        int r = ts.upload(ctx);
        if (r == B2G_OK)
            r = ts.run(ctx, psi, psi, dm.data);
        if (r != B2G_OK)
            return r;
        B2G_HIP(hipStreamSynchronize(ctx->stream));
        if (noise != 0.0) {
            B2G_REQUIRE(gen != nullptr, B2G_ERR_ARG, "noise without a generator");
            r = add_noise(ctx, wfn, psi_len, n_roots, dm, wfn_dm, ts, noise,
                          noise_type, *gen);
            if (r != B2G_OK)
                return r;
        }
    }
    tm.mark("density matrix");
    // ---- eigen-decomposition of every block
    double *E = nullptr;
    B2G_HIP(hipMalloc(&E, std::max<uint64_t>(dm.len, 1) * sizeof(double)));
    std::unique_ptr<double, int (*)(void *)> e_guard(
        E, [](void *p) { return (int)hipFree(p); });
    std::vector<double> w;
    {
        std::vector<uint32_t> n(dm.nb.begin(), dm.nb.end());
        const int r = syev_run(ctx, n, dm.shift, dm.data, E, w, nullptr);
        if (r != B2G_OK)
            return r;
    }
    tm.mark("eigensolver");
    h->eigvals = w;
    // ---- selection (truncate_density_matrix)
    std::vector<uint64_t> woff(nd + 1, 0);
    for (int b = 0; b < nd; b++)
        woff[b + 1] = woff[b] + dm.nb[b];
    std::vector<double> wr(w);
    for (int b = 0; b < nd; b++) {
        const double mult = (double)q_multiplicity(sy, dm.quanta[b]);
        for (uint64_t j = woff[b]; j < woff[b + 1]; j++) {
            if (trunc_type & 1u)
                wr[j] = w[j] * (1.0 / mult);
            else if (trunc_type & 2u)
                wr[j] = w[j] * mult;
        }
    }
    {
        const bool with_multi = (trunc_type & 8u) != 0;
        for (int b = 0; b < nd; b++) {
            const int reps = with_multi ? q_multiplicity(sy, dm.quanta[b]) : 1;
            for (int p = 0; p < reps; p++)
                for (uint64_t j = woff[b]; j < woff[b + 1]; j++)
                    h->spectra.push_back(std::sqrt(std::max(0.0, w[j])));
        }
    }
    std::vector<std::pair<int, int>> ss;
    for (int b = 0; b < nd; b++)
        for (uint32_t j = 0; j < dm.nb[b]; j++)
            ss.emplace_back(b, (int)j);
    const int k_total = (int)ss.size();
    auto wv = [&](const std::pair<int, int> &x) { return w[woff[x.first] + x.second]; };
    auto wrv = [&](const std::pair<int, int> &x) {
        return wr[woff[x.first] + x.second];
    };
    double error = 0.0;
    if (k != -1) {
        std::sort(ss.begin(), ss.end(),
                  [&](const std::pair<int, int> &a, const std::pair<int, int> &b) {
                      return wrv(a) > wrv(b);
                  });
        const unsigned keep = trunc_type / 16u;
        if (keep == 0) {
            for (int i = k; i < k_total; i++)
                if (wv(ss[i]) > 0)
                    error += wv(ss[i]);
            for (k = std::min(k, k_total); k > 1 && wrv(ss[k - 1]) < cutoff;
                 k--)
                if (wv(ss[k - 1]) > 0)
                    error += wv(ss[k - 1]);
            if (k < k_total)
                ss.resize(k);
        } else {
            std::vector<int> mask(nd, 0), smask(k_total, 0);
            for (int i = 0; i < k_total; i++) {
                mask[ss[i].first]++;
                smask[i] = mask[ss[i].first] > (int)keep;
            }
            for (int i = k; i < k_total; i++)
                if (wv(ss[i]) > 0 && smask[i])
                    error += wv(ss[i]);
            for (k = std::min(k, k_total); k > 1 && wrv(ss[k - 1]) < cutoff;
                 k--)
                if (wv(ss[k - 1]) > 0 && smask[k - 1])
                    error += wv(ss[k - 1]);
            std::vector<std::pair<int, int>> rss(
                ss.begin(), k < k_total ? ss.begin() + k : ss.end());
            for (int i = k; i < k_total; i++)
                if (!smask[i])
                    rss.push_back(ss[i]);
            ss.swap(rss);
        }
        std::sort(ss.begin(), ss.end());
    }
    h->error = error;
    // kept sectors and their state counts
    std::vector<int> ilr;
    std::vector<uint32_t> im;
    for (size_t i = 0; i < ss.size(); i++) {
        if (i == 0 || ss[i].first != ilr.back())
            ilr.push_back(ss[i].first), im.push_back(1);
        else
            im.back()++;
    }
    const int kk = (int)ilr.size();
    // ---- output infos (rotation_matrix_info_from_density_matrix,
    // wavefunction_info_from_density_matrix)
    b2g_split_s::Part &rot = h->part[trace_right ? 0 : 1];
    b2g_split_s::Part &wav = h->part[trace_right ? 1 : 0];
    rot.dq = dm.dq;
    for (int i = 0; i < kk; i++) {
        rot.quanta.push_back(dm.quanta[ilr[i]]);
        rot.nb.push_back(trace_right ? dm.nb[ilr[i]] : im[i]);
        rot.nk.push_back(trace_right ? im[i] : dm.nb[ilr[i]]);
    }
    set_shifts(rot);
    std::vector<std::vector<int>> dm_to_wfn(nd);
    for (int i = 0; i < nw; i++)
        dm_to_wfn[wfn_dm[i]].push_back(i);
    wav.dq = wdq;
    for (int i = 0; i < kk; i++)
        for (int iw : dm_to_wfn[ilr[i]]) {
            wav.quanta.push_back(wfn->quanta[iw]);
            wav.nb.push_back(trace_right ? im[i] : wfn->n_states_bra[iw]);
            wav.nk.push_back(trace_right ? wfn->n_states_ket[iw] : im[i]);
        }
    sort_part(wav);
    set_shifts(wav);
    B2G_HIP(hipMalloc(&rot.data, std::max<uint64_t>(rot.len, 1) * sizeof(double)));
    B2G_HIP(hipMalloc(&wav.data, std::max<uint64_t>(wav.len * n_roots, 1) *
                                     sizeof(double)));
    // ---- rotation matrix: kept eigenvectors (rows of E) as columns
    // (trace_right) or rows
    {
        std::vector<RowCopy> jobs;
        size_t iss = 0;
        for (int i = 0; i < kk; i++) {
            const int b = ilr[i];
            for (uint32_t j = 0; j < im[i]; j++, iss++) {
                RowCopy c;
                c.src = dm.shift[b] + (uint64_t)ss[iss].second * dm.nb[b];
                c.n = dm.nb[b];
                if (trace_right)
                    c.dst = rot.shift[i] + j, c.stride = im[i];
                else
                    c.dst = rot.shift[i] + (uint64_t)j * dm.nb[b], c.stride = 1;
                jobs.push_back(c);
            }
        }
        if (!jobs.empty()) {
            RowCopy *d_jobs = nullptr;
            B2G_HIP(hipMalloc(&d_jobs, jobs.size() * sizeof(RowCopy)));
            hipError_t e = hipMemcpyAsync(d_jobs, jobs.data(),
                                          jobs.size() * sizeof(RowCopy),
                                          hipMemcpyHostToDevice, ctx->stream);
            if (e == hipSuccess) {
                const unsigned g = (unsigned)std::min<size_t>(jobs.size(), 4096);
                hipLaunchKernelGGL(row_copy_kernel, dim3(g), dim3(256), 0,
                                   ctx->stream, d_jobs, (int)jobs.size(), E,
                                   rot.data);
                e = hipGetLastError();
            }
            const hipError_t e2 = hipStreamSynchronize(ctx->stream);
            hipFree(d_jobs);
            B2G_HIP(e);
            B2G_HIP(e2);
        }
    }
    // ---- the other factor: L^T psi (trace_right) or psi R^T, per root
    {
        TaskSet ts;
        for (int rt = 0; rt < n_roots; rt++)
        for (int i = 0; i < kk; i++)
            for (int iw : dm_to_wfn[ilr[i]]) {
                const int o = find_q(wav.quanta, wfn->quanta[iw]);
                B2G_REQUIRE(o >= 0, B2G_ERR_STATE, "split: block not found");
                Seg s;
                s.alpha = 1.0;
                uint32_t M, N;
                const uint64_t po = (uint64_t)rt * psi_len;
                if (trace_right) { // (im x n_bra)(n_bra x n_ket)
                    s.a_off = rot.shift[i], s.lda = im[i];
                    s.b_off = po + wfn->block_shifts[iw];
                    s.ldb = wfn->n_states_ket[iw];
                    s.K = dm.nb[ilr[i]], s.flags = 1;
                    M = im[i], N = wfn->n_states_ket[iw];
                } else { // (n_bra x n_ket)(n_ket x im)
                    s.a_off = po + wfn->block_shifts[iw];
                    s.lda = wfn->n_states_ket[iw];
                    s.b_off = rot.shift[i], s.ldb = dm.nb[ilr[i]];
                    s.K = dm.nb[ilr[i]], s.flags = 2;
                    M = wfn->n_states_bra[iw], N = im[i];
                }
                const uint32_t sb = (uint32_t)ts.s.size();
                if (s.K > 0)
                    ts.s.push_back(s);
                if (M > 0 && N > 0)
                    add_tiles(ts.t, (uint64_t)rt * wav.len + wav.shift[o], M,
                              N, s.K, sb, (uint32_t)ts.s.size(), 0);
            }
        int r = ts.upload(ctx);
        if (r == B2G_OK)
            r = trace_right ? ts.run(ctx, rot.data, psi, wav.data)
                            : ts.run(ctx, psi, rot.data, wav.data);
        if (r != B2G_OK)
            return r;
    }
    for (int rt = 0; normalize && wav.len > 0 && rt < n_roots; rt++) {
        double *wr = wav.data + (uint64_t)rt * wav.len;
        double nrm2 = 0.0;
        int r = b2g_vec_dot(ctx, wr, wr, wav.len, &nrm2);
        if (r != B2G_OK)
            return r;
        if (nrm2 > 0.0 &&
            (r = b2g_vec_scale(ctx, 1.0 / std::sqrt(nrm2), wr, wav.len)) !=
                B2G_OK)
            return r;
    }
    B2G_HIP(hipStreamSynchronize(ctx->stream));
    tm.mark("split");
    *out = h.release();
    return B2G_OK;
}

} // namespace b2g

extern "C" int b2g_split_density_matrix(b2g_ctx ctx, int symmetry,
                                        const b2g_sparse_info *wfn,
                                        const double *psi, int k,
                                        double cutoff, int trace_right,
                                        int normalize, unsigned trunc_type,
                                        double scale, b2g_split *out) {
    const double one = 1.0;
    return split_impl(ctx, symmetry, wfn, psi, 1, &one, k, cutoff,
                      trace_right, normalize, trunc_type, scale, 0.0, 0,
                      nullptr, out);
}

extern "C" int b2g_split_density_matrix_noise(
    b2g_ctx ctx, int symmetry, const b2g_sparse_info *wfn, const double *psi,
    int n_roots, const double *weights, double noise, int noise_type,
    unsigned seed, int k, double cutoff, int trace_right, int normalize,
    unsigned trunc_type, double scale, b2g_split *out) {
    // block2's Random::rand_seed(0) seeds from the clock, so no fixed stream
    // corresponds to seed 0; a caller that kept one default seed would also
    // add the same tensor at every site
    B2G_REQUIRE(noise == 0.0 || seed != 0u, B2G_ERR_ARG,
                "b2g_split_density_matrix_noise: seed 0 (block2 takes 0 as a "
                "clock seed); pass a nonzero seed, a new one per site");
    std::mt19937 gen(seed);
    return split_impl(ctx, symmetry, wfn, psi, n_roots, weights, k, cutoff,
                      trace_right, normalize, trunc_type, scale, noise,
                      noise_type, &gen, out);
}

extern "C" int b2g_rng_create(unsigned seed, b2g_rng *out) {
    B2G_REQUIRE(out != nullptr, B2G_ERR_ARG, "NULL arg");
    B2G_REQUIRE(seed != 0u, B2G_ERR_ARG,
                "b2g_rng_create: seed 0 (block2 takes 0 as a clock seed)");
    *out = new b2g_rng_s{std::mt19937(seed)};
    return B2G_OK;
}

extern "C" int b2g_rng_destroy(b2g_rng rng) {
    delete rng;
    return B2G_OK;
}

extern "C" int b2g_split_density_matrix_noise_rng(
    b2g_ctx ctx, int symmetry, const b2g_sparse_info *wfn, const double *psi,
    int n_roots, const double *weights, double noise, int noise_type,
    b2g_rng rng, int k, double cutoff, int trace_right, int normalize,
    unsigned trunc_type, double scale, b2g_split *out) {
    B2G_REQUIRE(noise == 0.0 || rng != nullptr, B2G_ERR_ARG,
                "b2g_split_density_matrix_noise_rng: NULL generator");
    return split_impl(ctx, symmetry, wfn, psi, n_roots, weights, k, cutoff,
                      trace_right, normalize, trunc_type, scale, noise,
                      noise_type, rng ? &rng->gen : nullptr, out);
}

extern "C" int b2g_split_density_matrix_multi(
    b2g_ctx ctx, int symmetry, const b2g_sparse_info *wfn, const double *psi,
    int n_roots, const double *weights, int k, double cutoff, int trace_right,
    int normalize, unsigned trunc_type, double scale, b2g_split *out) {
    return split_impl(ctx, symmetry, wfn, psi, n_roots, weights, k, cutoff,
                      trace_right, normalize, trunc_type, scale, 0.0, 0,
                      nullptr, out);
}

extern "C" int b2g_split_density_matrix_extra(
    b2g_ctx ctx, int symmetry, const b2g_sparse_info *wfn, const double *buf,
    int n_extra, const b2g_sparse_info *extra_infos,
    const uint64_t *extra_offsets, int k, double cutoff, int trace_right,
    int normalize, unsigned trunc_type, double scale, b2g_split *out) {
    B2G_REQUIRE(n_extra >= 0 && (n_extra == 0 || (extra_infos &&
                                                   extra_offsets)),
                B2G_ERR_ARG, "bad extra terms");
    ExtraTerms ex;
    ex.n = n_extra, ex.infos = extra_infos, ex.offs = extra_offsets;
    const double one = 1.0;
    return split_impl(ctx, symmetry, wfn, buf, 1, &one, k, cutoff,
                      trace_right, normalize, trunc_type, scale, 0.0, 0,
                      nullptr, out, ex);
}

extern "C" int b2g_split_n_roots(b2g_split s, int *n) {
    B2G_REQUIRE(s && n, B2G_ERR_ARG, "NULL arg");
    *n = s->n_roots;
    return B2G_OK;
}

extern "C" int b2g_split_result(b2g_split s, int which, b2g_sparse_info *info,
                                double **data, size_t *len) {
    B2G_REQUIRE(s && which >= 0 && which < 3, B2G_ERR_ARG, "bad argument");
    b2g_split_s::Part &p = s->part[which];
    if (info) {
        info->delta_quantum = p.dq;
        info->is_fermion = p.fermion;
        info->n = (int32_t)p.quanta.size();
        info->quanta = p.quanta.data();
        info->n_states_bra = p.nb.data();
        info->n_states_ket = p.nk.data();
        info->block_shifts = p.shift.data();
    }
    if (data)
        *data = p.data;
    if (len)
        *len = p.len;
    return B2G_OK;
}

extern "C" int b2g_split_error(b2g_split s, double *error) {
    B2G_REQUIRE(s && error, B2G_ERR_ARG, "NULL arg");
    *error = s->error;
    return B2G_OK;
}

extern "C" int b2g_split_values(b2g_split s, int which, double *out,
                                size_t *n) {
    B2G_REQUIRE(s && n && (which == 0 || which == 1), B2G_ERR_ARG,
                "bad argument");
    const std::vector<double> &v = which == 0 ? s->eigvals : s->spectra;
    if (out != nullptr) {
        B2G_REQUIRE(*n >= v.size(), B2G_ERR_ARG, "output array too short");
        std::copy(v.begin(), v.end(), out);
    }
    *n = v.size();
    return B2G_OK;
}

extern "C" int b2g_split_destroy(b2g_split s) {
    if (s == nullptr)
        return B2G_OK;
    if (s->ctx)
        hipStreamSynchronize(s->ctx->stream);
    delete s;
    return B2G_OK;
}
