// Device kernels: segmented grouped GEMM on MFMA (FP64 / FP32), the
// diagonal, and the FP64 vector algebra of the Davidson loop.
#include "kernels.hpp"
#include "launch.hpp"


namespace b2g {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename TC> struct Mfma;
// 16x16x4 single-block MFMA.  Lane l holds A[l%16][l/16] and B[l/16][l%16]
// for both types; accumulator element r of lane l is D[row][l%16] with
// row = 4*r + l/16 for FP64 and row = 4*(l/16) + r for FP32 (measured with
// an in-bounds probe of the builtins on gfx950).
template <> struct Mfma<double> {
    typedef v4d acc_t;
    static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) {
        return 4 * r + lk;
    }
};
template <> struct Mfma<float> {
    typedef v4f acc_t;
    static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) {
        return 4 * lk + r;
    }
};

// Raw buffer resource over [p, p + 2^31 - 16) bytes: loads at byte offsets
// at or past the record count return zero without touching memory (buffer
// range checking), which masks out-of-block elements.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0,
                                             0x7FFFFFF0, 0x00020000);
}
constexpr uint32_t kMasked = 0x80000000u; // offset past the record count

// buffer load at byte offset off + soff (off past the record count: zero)
template <typename T>
__device__ __forceinline__ T buf_load_s(__amdgpu_buffer_rsrc_t r, uint32_t off,
                                        uint32_t soff);
template <>
__device__ __forceinline__ double buf_load_s<double>(__amdgpu_buffer_rsrc_t r,
                                                     uint32_t off,
                                                     uint32_t soff) {
    return __builtin_bit_cast(
        double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, (int)soff, 0));
}
template <>
__device__ __forceinline__ float buf_load_s<float>(__amdgpu_buffer_rsrc_t r,
                                                   uint32_t off,
                                                   uint32_t soff) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, (int)soff, 0));
}

// One workgroup = 4 waves in a 2x2 arrangement; the workgroup tile is
// TM x TN, each wave owns (TM/2) x (TN/2) as (TM/32) x (TN/32) MFMA tiles.
// K is staged through two LDS buffers in chunks of TK: while the MFMAs run on
// one buffer, the next chunk's global loads are in flight in registers and
// are then written to the other buffer, so one barrier per chunk suffices.
// Element q of thread tid is e = tid + 256 q; since 256 is a multiple of TK,
// TM and TN, one coordinate of e is the same for all q and the other
// advances by a constant.  Tasks are strided over the grid (persistent
// launch); blockIdx.y selects the right-hand side.
//
// The registers of one chunk and their store to LDS ([k][m] and [k][n]
// layouts, one padding column), the per-segment alpha applied there.
template <typename TA, typename TB, typename TC, int TM, int TN, int TK>
struct TileRegs {
    static constexpr int NA = TM * TK / 256, NB = TN * TK / 256;
    static_assert(256 % TK == 0 && 256 % TM == 0 && 256 % TN == 0,
                  "tile shape");
    TA ra[NA];
    TB rb[NB];
    // registers -> LDS buffer ([k][m] and [k][n] layouts), alpha applied;
    // per layout, element q sits at a per-thread base plus a constant
    __device__ __forceinline__ void store(TC *As, TC *Bs, const Seg &sg,
                                          int tid) {
        const bool ta = sg.flags & 1, tb = (sg.flags >> 1) & 1;
        const TC alpha = (TC)sg.alpha;
        if (!ta) { // m = tid / TK + q * 256 / TK, k = tid % TK
            TC *p = As + (tid % TK) * (TM + 1) + tid / TK;
#pragma unroll
            for (int q = 0; q < NA; q++)
                p[q * (256 / TK)] = (TC)ra[q] * alpha;
        } else { // m = tid % TM, k = tid / TM + q * 256 / TM
            TC *p = As + (tid / TM) * (TM + 1) + tid % TM;
#pragma unroll
            for (int q = 0; q < NA; q++)
                p[q * (256 / TM) * (TM + 1)] = (TC)ra[q] * alpha;
        }
        if (!tb) { // n = tid % TN, k = tid / TN + q * 256 / TN
            TC *p = Bs + (tid / TN) * (TN + 1) + tid % TN;
#pragma unroll
            for (int q = 0; q < NB; q++)
                p[q * (256 / TN) * (TN + 1)] = (TC)rb[q];
        } else { // n = tid / TK + q * 256 / TK, k = tid % TK
            TC *p = Bs + (tid % TK) * (TN + 1) + tid / TK;
#pragma unroll
            for (int q = 0; q < NB; q++)
                p[q * (256 / TK)] = (TC)rb[q];
        }
    }
};

// Buffer loads at 32-bit byte offsets from the segment's blocks, with
// segment-invariant offsets: at the first chunk of a segment every
// element's byte offset at k0 = 0 is computed once (rows or columns outside
// the block get an offset past the buffer's record count and read as zero);
// each chunk then passes its k0 as the instruction's scalar offset, so a
// full chunk costs no vector arithmetic before its loads.  Elements at
// k >= K (the last, partial chunk of a segment) get the masked offset per
// element.  Measured on the n=40 M=3000 middle site against 64-bit
// addresses recomputed per chunk: FP64 H*psi 54.3 -> 48.0 ms per 4-RHS step,
// its 128 x 128 phase-2 launch 17.5 -> 14.9 ms.
template <typename TA, typename TB, typename TC, int TM, int TN, int TK>
struct TileLoaderS : TileRegs<TA, TB, TC, TM, TN, TK> {
    using Base = TileRegs<TA, TB, TC, TM, TN, TK>;
    static constexpr int NA = Base::NA, NB = Base::NB;
    uint32_t va[NA], vb[NB]; // byte offsets at k0 = 0 (kMasked: outside)
    uint32_t sa, sb;         // bytes per unit of k0
    __amdgpu_buffer_rsrc_t rA, rB;
    __device__ __forceinline__ void setup(const TA *__restrict__ A,
                                          const TB *__restrict__ B,
                                          const Seg &sg, const Task &task,
                                          int tid) {
        const bool ta = sg.flags & 1, tb = (sg.flags >> 1) & 1;
        rA = buf_rsrc(A + sg.a_off);
        rB = buf_rsrc(B + sg.b_off);
        if (!ta) { // (m, k) at m * lda + k; k = tid % TK
            const uint32_t k = tid % TK;
            uint32_t r = task.row0 + tid / TK;
#pragma unroll
            for (int q = 0; q < NA; q++) {
                va[q] = r < task.M ? (r * sg.lda + k) * (uint32_t)sizeof(TA)
                                   : kMasked;
                r += 256 / TK;
            }
            sa = (uint32_t)sizeof(TA);
        } else { // (m, k) at k * lda + m; m = tid % TM
            const uint32_t m = task.row0 + tid % TM;
            uint32_t k = tid / TM;
#pragma unroll
            for (int q = 0; q < NA; q++) {
                va[q] = m < task.M ? (k * sg.lda + m) * (uint32_t)sizeof(TA)
                                   : kMasked;
                k += 256 / TM;
            }
            sa = sg.lda * (uint32_t)sizeof(TA);
        }
        if (!tb) { // (k, n) at k * ldb + n; n = tid % TN
            const uint32_t n = task.col0 + tid % TN;
            uint32_t k = tid / TN;
#pragma unroll
            for (int q = 0; q < NB; q++) {
                vb[q] = n < task.N ? (k * sg.ldb + n) * (uint32_t)sizeof(TB)
                                   : kMasked;
                k += 256 / TN;
            }
            sb = sg.ldb * (uint32_t)sizeof(TB);
        } else { // (k, n) at n * ldb + k; k = tid % TK
            const uint32_t k = tid % TK;
            uint32_t c = task.col0 + tid / TK;
#pragma unroll
            for (int q = 0; q < NB; q++) {
                vb[q] = c < task.N ? (c * sg.ldb + k) * (uint32_t)sizeof(TB)
                                   : kMasked;
                c += 256 / TK;
            }
            sb = (uint32_t)sizeof(TB);
        }
        sa = __builtin_amdgcn_readfirstlane(sa);
        sb = __builtin_amdgcn_readfirstlane(sb);
    }
    __device__ __forceinline__ void load(const TA *__restrict__ A,
                                         const TB *__restrict__ B,
                                         const Seg &sg, uint32_t k0,
                                         const Task &task, int tid) {
        if (k0 == 0)
            setup(A, B, sg, task, tid);
        const uint32_t soa = __builtin_amdgcn_readfirstlane(k0 * sa);
        const uint32_t sob = __builtin_amdgcn_readfirstlane(k0 * sb);
        if (k0 + TK <= sg.K) {
#pragma unroll
            for (int q = 0; q < NA; q++)
                this->ra[q] = buf_load_s<TA>(rA, va[q], soa);
#pragma unroll
            for (int q = 0; q < NB; q++)
                this->rb[q] = buf_load_s<TB>(rB, vb[q], sob);
            return;
        }
        // partial chunk: the k of element q, per storage order
        const bool ta = sg.flags & 1, tb = (sg.flags >> 1) & 1;
#pragma unroll
        for (int q = 0; q < NA; q++) {
            const uint32_t k =
                k0 + (ta ? tid / TM + q * (256 / TM) : tid % TK);
            this->ra[q] = buf_load_s<TA>(rA, k < sg.K ? va[q] : kMasked, soa);
        }
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const uint32_t k =
                k0 + (tb ? tid % TK : tid / TN + q * (256 / TN));
            this->rb[q] = buf_load_s<TB>(rB, k < sg.K ? vb[q] : kMasked, sob);
        }
    }
};

template <typename TA, typename TB, typename TD, typename TC, int TM, int TN,
          int TK, int PHASE>
__global__ __launch_bounds__(256) void seg_gemm_kernel(
    const TA *__restrict__ A, const TB *__restrict__ B, TD *__restrict__ D,
    TD *__restrict__ P, const Task *__restrict__ tasks, int ntasks,
    const Seg *__restrict__ segs, uint64_t a_rs, uint64_t b_rs, uint64_t d_rs,
    uint64_t p_rs, double scale, int beta0, XcdSplit xs) {
    constexpr int WM = TM / 2, WN = TN / 2;
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int ASZ = TK * (TM + 1), BSZ = TK * (TN + 1);
    typedef typename Mfma<TC>::acc_t acc_t;
    __shared__ TC smem[2 * (ASZ + BSZ)];
    const int rhs = blockIdx.y;
    A += rhs * a_rs;
    B += rhs * b_rs;
    D += rhs * d_rs;
    P += rhs * p_rs;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int lr = lane & 15, lk = lane >> 4;
    TileLoaderS<TA, TB, TC, TM, TN, TK> ld;
    int t0 = blockIdx.x, tend = ntasks, tstep = gridDim.x;
    if (xs.off[8] != 0) {
        const int x = blockIdx.x & 7;
        t0 = (int)xs.off[x] + (int)(blockIdx.x >> 3);
        tend = (int)xs.off[x + 1];
        tstep = gridDim.x >> 3;
    }
    for (int t = t0; t < tend; t += tstep) {
        const Task task = tasks[t];
        acc_t acc[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FN; j++)
                acc[i][j] = acc_t{0, 0, 0, 0};
        // segments have K > 0 (the plan drops empty ones); descriptors are
        // wave-uniform and read one segment ahead
        uint32_t s = task.seg_begin, k0 = 0;
        if (s < task.seg_end) {
            Seg cur = segs[s];
            Seg nxt = cur;
            if (s + 1 < task.seg_end)
                nxt = segs[s + 1];
            int buf = 0;
            ld.load(A, B, cur, k0, task, tid);
            ld.store(smem, smem + ASZ, cur, tid);
            __syncthreads();
            while (true) {
                // advance the loader to the next chunk and start its loads
                k0 += TK;
                if (k0 >= cur.K) {
                    k0 = 0;
                    s++;
                    cur = nxt;
                    if (s + 1 < task.seg_end)
                        nxt = segs[s + 1];
                }
                const bool more = s < task.seg_end;
                if (more)
                    ld.load(A, B, cur, k0, task, tid);
                const TC *As = smem + buf * (ASZ + BSZ);
                const TC *Bs = As + ASZ;
                // MFMAs on this buffer, fragment reads one k-step ahead
                TC a[2][FM], b[2][FN];
#pragma unroll
                for (int i = 0; i < FM; i++)
                    a[0][i] = As[lk * (TM + 1) + wm * WM + i * 16 + lr];
#pragma unroll
                for (int j = 0; j < FN; j++)
                    b[0][j] = Bs[lk * (TN + 1) + wn * WN + j * 16 + lr];
#pragma unroll
                for (int ks = 0; ks < TK / 4; ks++) {
                    const int cb = ks & 1, nb = cb ^ 1;
                    if (ks + 1 < TK / 4) {
                        const int kr = (ks + 1) * 4 + lk;
#pragma unroll
                        for (int i = 0; i < FM; i++)
                            a[nb][i] = As[kr * (TM + 1) + wm * WM + i * 16 + lr];
#pragma unroll
                        for (int j = 0; j < FN; j++)
                            b[nb][j] = Bs[kr * (TN + 1) + wn * WN + j * 16 + lr];
                    }
#pragma unroll
                    for (int i = 0; i < FM; i++)
#pragma unroll
                        for (int j = 0; j < FN; j++)
                            acc[i][j] = Mfma<TC>::op(a[cb][i], b[cb][j],
                                                     acc[i][j]);
                }
                if (!more)
                    break;
                // the other buffer was released by the barrier that ended
                // the previous chunk
                buf ^= 1;
                TC *Ns = smem + buf * (ASZ + BSZ);
                ld.store(Ns, Ns + ASZ, cur, tid);
                __syncthreads();
            }
            __syncthreads(); // all waves done with LDS before the next task
        }
        mfma_settle<acc_t, FM * FN>(&acc[0][0]);
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FN; j++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const uint32_t row =
                        task.row0 + wm * WM + i * 16 + Mfma<TC>::row(lk, r);
                    const uint32_t col = task.col0 + wn * WN + j * 16 + lr;
                    if (row < task.M && col < task.N) {
                        const uint64_t idx =
                            task.d_off + (uint64_t)row * task.ldd + col;
                        const double v = scale * (double)acc[i][j][r];
                        if (task.beta & 2)
                            P[idx] = (TD)v; // partial tile, reduced later
                        else if ((task.beta | beta0) & 1)
                            D[idx] = (TD)((double)D[idx] + v);
                        else
                            D[idx] = (TD)v;
                    }
                }
    }
}

// One workgroup reduces a 1024-element slice of one tile (tiles hold at most
// 128 x 128 = 16 slices): thread i takes elements 4i..4i+3 of the slice,
// reading each part's four values as two 16-byte loads (partial slots are
// 1024-element aligned and tile widths multiples of 32, so the four share a
// row).  Parts are summed in increasing order: the result does not depend on
// the launch shape.
constexpr int kRedSlices = 16;

template <typename TD>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(
    TD *__restrict__ D, const TD *__restrict__ P,
    const RedTask *__restrict__ tasks, int ntasks, uint64_t d_rs,
    uint64_t p_rs, int beta0) {
    const int rhs = blockIdx.y;
    D += rhs * d_rs;
    P += rhs * p_rs;
    const int64_t nwork = (int64_t)ntasks * kRedSlices;
    for (int64_t w = blockIdx.x; w < nwork; w += gridDim.x) {
        const RedTask rt = tasks[w / kRedSlices];
        const uint32_t tsz = rt.tm * rt.tn;
        const uint32_t e = (uint32_t)(w % kRedSlices) * 1024 + 4 * threadIdx.x;
        if (e >= tsz)
            continue;
        const uint32_t r = e / rt.tn, c = e % rt.tn;
        const uint32_t row = rt.row0 + r, col = rt.col0 + c;
        if (row >= rt.M || col >= rt.N)
            continue;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        const TD *q = P + rt.p_off + e;
#pragma unroll 4
        for (uint32_t p = 0; p < rt.nparts; p++) {
            const TD *qp = q + (uint64_t)p * tsz;
            s0 += (double)qp[0];
            s1 += (double)qp[1];
            s2 += (double)qp[2];
            s3 += (double)qp[3];
        }
        const double sv[4] = {s0, s1, s2, s3};
        const uint64_t idx = rt.d_off + (uint64_t)row * rt.ldd + col;
        const bool acc = (rt.beta | beta0) & 1;
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (col + i < rt.N)
                D[idx + i] = acc ? (TD)((double)D[idx + i] + sv[i]) : (TD)sv[i];
    }
}

template <typename TD>
hipError_t launch_splitk_reduce(TD *D, const TD *P, const RedTask *tasks,
                                int ntasks, uint64_t d_rs, uint64_t p_rs,
                                int beta0, int n_rhs, int grid,
                                hipStream_t st) {
    if (ntasks <= 0)
        return hipSuccess;
    const int64_t nwork = (int64_t)ntasks * kRedSlices;
    const int64_t cap = (int64_t)grid * 4;
    dim3 g((unsigned)(nwork < cap ? nwork : cap), (unsigned)n_rhs);
    hipLaunchKernelGGL(splitk_reduce_kernel<TD>, g, dim3(256), 0, st, D, P,
                       tasks, ntasks, d_rs, p_rs, beta0);
    return hipGetLastError();
}
template hipError_t launch_splitk_reduce<double>(double *, const double *,
                                                 const RedTask *, int,
                                                 uint64_t, uint64_t, int, int,
                                                 int, hipStream_t);
template hipError_t launch_splitk_reduce<float>(float *, const float *,
                                                const RedTask *, int, uint64_t,
                                                uint64_t, int, int, int,
                                                hipStream_t);

// tile configurations (index = Task bucket id): cfg = 3 * i + j with
// TM = 32 << i and TN = 32 << j; K chunk 16 (32 for the 32x32 tile)
int tile_m(int cfg) { return 32 << (cfg / 3); }
int tile_n(int cfg) { return 32 << (cfg % 3); }

// persistent grid: min(grid, ntasks) workgroups
unsigned task_grid(int grid, int ntasks, const XcdSplit &xs) {
    const int g = grid < ntasks ? grid : ntasks;
    return (unsigned)(xs.off[8] != 0 ? (g + 7) & ~7 : g);
}

template <typename TA, typename TB, typename TD, typename TC, int PHASE>
hipError_t launch_seg_gemm(int cfg, const TA *A, const TB *B, TD *D, TD *P,
                           const Task *tasks, int ntasks, const Seg *segs,
                           uint64_t a_rs, uint64_t b_rs, uint64_t d_rs,
                           uint64_t p_rs, double scale, int beta0, int n_rhs,
                           int grid, hipStream_t st, XcdSplit xs) {
    if (ntasks <= 0)
        return hipSuccess;
    dim3 g(task_grid(grid, ntasks, xs), (unsigned)n_rhs);
#define B2G_LAUNCH(TM_, TN_, TK_)                                              \
    hipLaunchKernelGGL((seg_gemm_kernel<TA, TB, TD, TC, TM_, TN_, TK_, PHASE>), \
                       g,                                                      \
                       dim3(256), 0, st, A, B, D, P, tasks, ntasks, segs,      \
                       a_rs, b_rs, d_rs, p_rs, scale, beta0, xs)
    switch (cfg) {
    case 0: B2G_LAUNCH(32, 32, 32); break;
    case 1: B2G_LAUNCH(32, 64, 16); break;
    case 2: B2G_LAUNCH(32, 128, 16); break;
    case 3: B2G_LAUNCH(64, 32, 16); break;
    case 4: B2G_LAUNCH(64, 64, 16); break;
    case 5: B2G_LAUNCH(64, 128, 16); break;
    case 6: B2G_LAUNCH(128, 32, 16); break;
    case 7: B2G_LAUNCH(128, 64, 16); break;
    default: B2G_LAUNCH(128, 128, 16); break;
    }
#undef B2G_LAUNCH
    return hipGetLastError();
}

template hipError_t launch_seg_gemm<double, double, double, double, 1>(
    int, const double *, const double *, double *, double *, const Task *, int, const Seg *,
    uint64_t, uint64_t, uint64_t, uint64_t, double, int, int, int,
    hipStream_t, XcdSplit);
template hipError_t launch_seg_gemm<double, double, double, double, 2>(
    int, const double *, const double *, double *, double *, const Task *, int, const Seg *,
    uint64_t, uint64_t, uint64_t, uint64_t, double, int, int, int,
    hipStream_t, XcdSplit);
template hipError_t launch_seg_gemm<double, double, double, double, 3>(
    int, const double *, const double *, double *, double *, const Task *, int, const Seg *,
    uint64_t, uint64_t, uint64_t, uint64_t, double, int, int, int,
    hipStream_t, XcdSplit);
template hipError_t launch_seg_gemm<double, double, double, double, 4>(
    int, const double *, const double *, double *, double *, const Task *, int, const Seg *,
    uint64_t, uint64_t, uint64_t, uint64_t, double, int, int, int,
    hipStream_t, XcdSplit);
template hipError_t launch_seg_gemm<double, double, double, double, 5>(
    int, const double *, const double *, double *, double *, const Task *, int, const Seg *,
    uint64_t, uint64_t, uint64_t, uint64_t, double, int, int, int,
    hipStream_t, XcdSplit);
// 6: block-product lists of environment moves (two-site wavefunctions)
template hipError_t launch_seg_gemm<double, double, double, double, 6>(
    int, const double *, const double *, double *, double *, const Task *, int, const Seg *,
    uint64_t, uint64_t, uint64_t, uint64_t, double, int, int, int,
    hipStream_t, XcdSplit);
template hipError_t launch_seg_gemm<double, float, float, float, 1>(
    int, const double *, const float *, float *, float *, const Task *, int, const Seg *,
    uint64_t, uint64_t, uint64_t, uint64_t, double, int, int, int,
    hipStream_t, XcdSplit);
template hipError_t launch_seg_gemm<float, float, double, float, 2>(
    int, const float *, const float *, double *, double *, const Task *, int, const Seg *,
    uint64_t, uint64_t, uint64_t, uint64_t, double, int, int, int,
    hipStream_t, XcdSplit);

// ---------------------------------------------------------------- diagonal
__global__ __launch_bounds__(256) void diag_kernel(
    const double *__restrict__ L, const double *__restrict__ R,
    double *__restrict__ Dv, const DiagBlock *__restrict__ blocks, int nblocks,
    const DiagTermDev *__restrict__ terms) {
    for (int bi = blockIdx.x; bi < nblocks; bi += gridDim.x) {
        const DiagBlock blk = blocks[bi];
        const uint32_t mn = blk.m * blk.n;
        for (uint32_t e = threadIdx.x; e < mn; e += blockDim.x) {
            const uint32_t i = blk.row0 + e / blk.n, j = e % blk.n;
            double s = 0.0;
            for (uint32_t t = blk.t_begin; t < blk.t_end; t++) {
                const DiagTermDev tm = terms[t];
                // the term's a_n x b_n window starts at (r0, c0)
                const uint32_t ii = i - tm.r0, jj = j - tm.c0;
                if (ii < tm.a_n && jj < tm.b_n)
                    s += tm.alpha * L[tm.a_off + (uint64_t)ii * (tm.a_n + 1)] *
                         R[tm.b_off + (uint64_t)jj * (tm.b_n + 1)];
            }
            Dv[blk.d_off + e] = s;
        }
    }
}

hipError_t launch_diag(const double *L, const double *R, double *Dv,
                       const DiagBlock *blocks, int nblocks,
                       const DiagTermDev *terms, int grid, hipStream_t st) {
    if (nblocks <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(diag_kernel, dim3(grid < nblocks ? grid : nblocks),
                       dim3(256), 0, st, L, R, Dv, blocks, nblocks, terms);
    return hipGetLastError();
}

// ------------------------------------------------------------ conversions
template <typename TS, typename TT>
__global__ void convert_kernel(const TS *__restrict__ s, TT *__restrict__ t,
                               size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        t[i] = (TT)s[i];
}
hipError_t launch_f64_to_f32(const double *s, float *t, size_t n,
                             hipStream_t st) {
    if (n == 0)
        return hipSuccess;
    size_t g = (n + 255) / 256;
    if (g > 65536)
        g = 65536;
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3((unsigned)g),
                       dim3(256), 0, st, s, t, n);
    return hipGetLastError();
}

// ------------------------------------------------------------ vector ops
// Two-stage deterministic reductions: stage 1 writes one partial per block
// (fixed grid), stage 2 sums the partials in order in a single block.
constexpr int kRedGrid = 512;

__device__ __forceinline__ double block_sum(double v, double *sh) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s)
            sh[tid] += sh[tid + s];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// partial[b*k + i] = sum over the block's rows of X_i . Y_i  (k pairs)
__global__ __launch_bounds__(256) void dot_stage1(
    const double *__restrict__ X, const double *__restrict__ Y, size_t n,
    double *__restrict__ partial) {
    __shared__ double sh[256];
    double s = 0.0;
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n;
         i += (size_t)gridDim.x * 256)
        s += X[i] * Y[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0)
        partial[blockIdx.x] = s;
}

// out[i] = sum_b partial[b * stride + i] for i < k
__global__ __launch_bounds__(256) void reduce_stage2(
    const double *__restrict__ partial, int nblk, int k, double *__restrict__ out) {
    __shared__ double sh[256];
    for (int i = 0; i < k; i++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblk; b += 256)
            s += partial[(size_t)b * k + i];
        s = block_sum(s, sh);
        if (threadIdx.x == 0)
            out[i] = s;
    }
}

hipError_t launch_dot(const double *x, const double *y, size_t n,
                      double *partial, double *out, hipStream_t st) {
    hipLaunchKernelGGL(dot_stage1, dim3(kRedGrid), dim3(256), 0, st, x, y, n,
                       partial);
    hipLaunchKernelGGL(reduce_stage2, dim3(1), dim3(256), 0, st, partial,
                       kRedGrid, 1, out);
    return hipGetLastError();
}

// Gram block  G[i*ky + j] = X_i . Y_j  for i < kx, j < ky <= 64 (vectors
// stored with stride ld).  Rows are processed in tiles of 32 staged through
// LDS.
__global__ __launch_bounds__(256) void gram_stage1(
    const double *__restrict__ X, int kx, const double *__restrict__ Y, int ky,
    size_t ld, size_t n, double *__restrict__ partial) {
    constexpr int RT = 32;
    __shared__ double xs[64][RT + 1];
    __shared__ double ys[64][RT + 1];
    const int npair = kx * ky;
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; q++)
        acc[q] = 0.0;
    for (size_t r0 = (size_t)blockIdx.x * RT; r0 < n;
         r0 += (size_t)gridDim.x * RT) {
        for (int e = threadIdx.x; e < kx * RT; e += 256) {
            const int i = e / RT, r = e % RT;
            xs[i][r] = r0 + r < n ? X[(size_t)i * ld + r0 + r] : 0.0;
        }
        for (int e = threadIdx.x; e < ky * RT; e += 256) {
            const int j = e / RT, r = e % RT;
            ys[j][r] = r0 + r < n ? Y[(size_t)j * ld + r0 + r] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int p = threadIdx.x + q * 256;
            if (p < npair) {
                const int i = p / ky, j = p % ky;
                double s = 0.0;
                for (int r = 0; r < RT; r++)
                    s += xs[i][r] * ys[j][r];
                acc[q] += s;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int p = threadIdx.x + q * 256;
        if (p < npair)
            partial[(size_t)blockIdx.x * npair + p] = acc[q];
    }
}

hipError_t launch_gram(const double *X, int kx, const double *Y, int ky,
                       size_t ld, size_t n, double *partial, double *out,
                       hipStream_t st) {
    if (kx <= 0 || ky <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(gram_stage1, dim3(kRedGrid), dim3(256), 0, st, X, kx,
                       Y, ky, ld, n, partial);
    hipLaunchKernelGGL(reduce_stage2, dim3(1), dim3(256), 0, st, partial,
                       kRedGrid, kx * ky, out);
    return hipGetLastError();
}

__global__ void axpy_kernel(double a, const double *__restrict__ x,
                            double *__restrict__ y, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        y[i] += a * x[i];
}
// y += coef[0] * x with the coefficient read on the device
__global__ void axpy_dev_kernel(const double *__restrict__ coef, double sign,
                                const double *__restrict__ x,
                                double *__restrict__ y, size_t n) {
    const double a = sign * coef[0];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        y[i] += a * x[i];
}
__global__ void scale_kernel(double a, double *__restrict__ x, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        x[i] *= a;
}
__global__ void copy_kernel(const double *__restrict__ x,
                            double *__restrict__ y, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        y[i] = x[i];
}
__global__ void fill_kernel(double v, double *__restrict__ x, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        x[i] = v;
}

static dim3 grid_for(size_t n) {
    size_t g = (n + 255) / 256;
    if (g > 8192)
        g = 8192;
    if (g == 0)
        g = 1;
    return dim3((unsigned)g);
}

hipError_t launch_axpy(double a, const double *x, double *y, size_t n,
                       hipStream_t st) {
    hipLaunchKernelGGL(axpy_kernel, grid_for(n), dim3(256), 0, st, a, x, y, n);
    return hipGetLastError();
}
hipError_t launch_axpy_dev(const double *coef, double sign, const double *x,
                           double *y, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(axpy_dev_kernel, grid_for(n), dim3(256), 0, st, coef,
                       sign, x, y, n);
    return hipGetLastError();
}
hipError_t launch_scale(double a, double *x, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(scale_kernel, grid_for(n), dim3(256), 0, st, a, x, n);
    return hipGetLastError();
}
hipError_t launch_copy(const double *x, double *y, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(copy_kernel, grid_for(n), dim3(256), 0, st, x, y, n);
    return hipGetLastError();
}
hipError_t launch_fill(double v, double *x, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(fill_kernel, grid_for(n), dim3(256), 0, st, v, x, n);
    return hipGetLastError();
}

// X_j <- sum_i V[j*m + i] X_i for j < m, in place; vectors with stride ld.
// Each thread owns one row across all m vectors.
template <int MAXM>
__global__ __launch_bounds__(256) void rotate_kernel(
    double *__restrict__ X, size_t ld, size_t n, int m,
    const double *__restrict__ V) {
    __shared__ double vs[MAXM * MAXM];
    for (int e = threadIdx.x; e < m * m; e += 256)
        vs[e] = V[e];
    __syncthreads();
    for (size_t r = blockIdx.x * (size_t)256 + threadIdx.x; r < n;
         r += (size_t)gridDim.x * 256) {
        double x[MAXM];
#pragma unroll
        for (int i = 0; i < MAXM; i++)
            x[i] = i < m ? X[(size_t)i * ld + r] : 0.0;
#pragma unroll
        for (int j = 0; j < MAXM; j++) {
            if (j < m) {
                double s = 0.0;
#pragma unroll
                for (int i = 0; i < MAXM; i++)
                    if (i < m)
                        s += vs[j * m + i] * x[i];
                X[(size_t)j * ld + r] = s;
            }
        }
    }
}

hipError_t launch_rotate(double *X, size_t ld, size_t n, int m,
                         const double *V, hipStream_t st) {
    dim3 g = grid_for(n);
    if (m <= 8)
        hipLaunchKernelGGL((rotate_kernel<8>), g, dim3(256), 0, st, X, ld, n,
                           m, V);
    else if (m <= 16)
        hipLaunchKernelGGL((rotate_kernel<16>), g, dim3(256), 0, st, X, ld, n,
                           m, V);
    else if (m <= 32)
        hipLaunchKernelGGL((rotate_kernel<32>), g, dim3(256), 0, st, X, ld, n,
                           m, V);
    else if (m <= 64)
        hipLaunchKernelGGL((rotate_kernel<64>), g, dim3(256), 0, st, X, ld, n,
                           m, V);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// Olsen preconditioner (IterativeMatrixFunctions::olsen_precondition,
// src/core/iterative_matrix_functions.hpp:93-108), first pass:
//   t = c / (ld - aa), q = q / (ld - aa) where |ld - aa| > 1e-12,
// plus partial sums of c.q and c.t (after scaling).
__global__ __launch_bounds__(256) void olsen_stage1(
    double *__restrict__ q, const double *__restrict__ c,
    const double *__restrict__ aa, double ld, double *__restrict__ t, size_t n,
    double *__restrict__ partial) {
    __shared__ double sh[256];
    double s_cq = 0.0, s_ct = 0.0;
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n;
         i += (size_t)gridDim.x * 256) {
        double ci = c[i], qi = q[i], ti = ci;
        const double dd = ld - aa[i];
        if (fabs(dd) > 1E-12) {
            ti = ci / dd;
            qi = qi / dd;
            q[i] = qi;
        }
        t[i] = ti;
        s_cq += ci * qi;
        s_ct += ci * ti;
    }
    s_cq = block_sum(s_cq, sh);
    s_ct = block_sum(s_ct, sh);
    if (threadIdx.x == 0) {
        partial[(size_t)blockIdx.x * 2 + 0] = s_cq;
        partial[(size_t)blockIdx.x * 2 + 1] = s_ct;
    }
}
// q += t * (-(c.q) / (c.t))
__global__ void olsen_stage2(double *__restrict__ q,
                             const double *__restrict__ t,
                             const double *__restrict__ sums, size_t n) {
    const double a = -sums[0] / sums[1];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        q[i] += a * t[i];
}

hipError_t launch_olsen(double *q, const double *c, const double *aa,
                        double ld, double *t, size_t n, double *partial,
                        double *sums, hipStream_t st) {
    hipLaunchKernelGGL(olsen_stage1, dim3(kRedGrid), dim3(256), 0, st, q, c,
                       aa, ld, t, n, partial);
    hipLaunchKernelGGL(reduce_stage2, dim3(1), dim3(256), 0, st, partial,
                       kRedGrid, 2, sums);
    hipLaunchKernelGGL(olsen_stage2, grid_for(n), dim3(256), 0, st, q, t, sums,
                       n);
    return hipGetLastError();
}

// Davidson's diagonal preconditioner (IterativeMatrixFunctions::
// davidson_precondition, src/core/iterative_matrix_functions.hpp:66-72):
// q = q / (ld - aa) where |ld - aa| > 1e-12
__global__ __launch_bounds__(256) void dav_precond_kernel(
    double *__restrict__ q, const double *__restrict__ aa, double ld,
    size_t n) {
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n;
         i += (size_t)gridDim.x * 256) {
        const double d = ld - aa[i];
        if (fabs(d) > 1E-12)
            q[i] /= d;
    }
}
hipError_t launch_dav_precond(double *q, const double *aa, double ld,
                              size_t n, hipStream_t st) {
    hipLaunchKernelGGL(dav_precond_kernel, grid_for(n), dim3(256), 0, st, q,
                       aa, ld, n);
    return hipGetLastError();
}

int red_grid() { return kRedGrid; }

} // namespace b2g
