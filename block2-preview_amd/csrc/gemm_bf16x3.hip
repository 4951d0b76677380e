// Segmented grouped GEMM in split-BF16 three-term arithmetic (gfx950).
//
// Every operand value a (FP64 or FP32 in memory) is rounded to FP32 and split
// into two BF16 numbers, hi = bf16(a) and lo = bf16(a - hi), so that
// a = hi + lo to about 2^-17 relative.  A product is then
//     a * b ~= a_hi * b_hi + a_hi * b_lo + a_lo * b_hi
// (the dropped a_lo * b_lo is ~2^-16 relative), each term one
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  Three BF16 MFMAs cost
// 3 * 16 cycles per 16x16x32 block against 8 * 32 cycles for the same block
// with FP32 16x16x4 MFMAs, so the split trades a ~1e-5 relative error for
// 5x the matrix throughput; the Davidson that uses it refines in FP64.
//
// Task / segment semantics are those of seg_gemm_kernel (kernels.hip): a
// task is one TM x TN output tile accumulating sum_s alpha_s op(A_s) op(B_s)
// over its segments, the tile owned by exactly one task (no atomics).
#include "kernels.hpp"
#include "launch.hpp"

namespace b2g {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Lane l of v_mfma_f32_16x16x32_bf16 supplies A[l%16][8*(l/16) + i] and
// B[8*(l/16) + i][l%16] (i < 8); accumulator element r of lane l is
// D[4*(l/16) + r][l%16] (in-bounds probe of the builtin on gfx950).
__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void split4(const float v[4], bf16x4 &hi,
                                       bf16x4 &lo) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const __bf16 h = (__bf16)v[i];
        hi[i] = h;
        lo[i] = (__bf16)(v[i] - (float)h);
    }
}

// One workgroup = 4 waves (2x2); K staged through two LDS buffers in chunks
// of 32 (one MFMA k-step).  LDS rows hold 32 k values of one tile row (A) or
// one tile column (B) plus 8 of padding, so that a lane's 8 k values are one
// 16-byte read and the 16 rows of a fragment fall on distinct banks.
template <typename TA, typename TB, int TM, int TN> struct Bf3Tile {
    static constexpr int TK = 32, KG = TK / 4, LDK = TK + 8;
    static constexpr int GA = TM * KG / 256, GB = TN * KG / 256;
    static_assert(GA >= 1 && GB >= 1 && 4 * (GA + GB) <= 32, "tile shape");
    // per-buffer element counts (BF16): A hi, A lo, B hi, B lo
    static constexpr int ASZ = TM * LDK, BSZ = TN * LDK;
    static constexpr int BUF = 2 * ASZ + 2 * BSZ;
    TA ra[GA][4];
    TB rb[GB][4];
    uint32_t okmask;

    // (row, k-group) of load group g of thread tid: consecutive lanes run
    // along the dimension that is contiguous in memory
    static __device__ __forceinline__ void a_coord(bool ta, int e, int &m,
                                                   int &kg) {
        m = ta ? e % TM : e / KG;
        kg = ta ? e / TM : e % KG;
    }
    static __device__ __forceinline__ void b_coord(bool tb, int e, int &n,
                                                   int &kg) {
        n = tb ? e / KG : e % TN;
        kg = tb ? e % KG : e / TN;
    }

    __device__ __forceinline__ void load(const TA *__restrict__ A,
                                         const TB *__restrict__ B,
                                         const Seg &sg, uint32_t k0,
                                         const Task &task, int tid) {
        const bool ta = sg.flags & 1, tb = (sg.flags >> 1) & 1;
        uint32_t mask = 0;
#pragma unroll
        for (int g = 0; g < GA; g++) {
            int m, kg;
            a_coord(ta, tid + g * 256, m, kg);
            const uint32_t r = task.row0 + m;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint32_t kk = k0 + 4 * kg + i;
                const bool ok = r < task.M && kk < sg.K;
                const uint64_t idx =
                    ok ? sg.a_off + (ta ? (uint64_t)kk * sg.lda + r
                                        : (uint64_t)r * sg.lda + kk)
                       : sg.a_off;
                ra[g][i] = A[idx];
                mask |= (uint32_t)ok << (4 * g + i);
            }
        }
#pragma unroll
        for (int g = 0; g < GB; g++) {
            int n, kg;
            b_coord(tb, tid + g * 256, n, kg);
            const uint32_t c = task.col0 + n;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint32_t kk = k0 + 4 * kg + i;
                const bool ok = c < task.N && kk < sg.K;
                const uint64_t idx =
                    ok ? sg.b_off + (tb ? (uint64_t)c * sg.ldb + kk
                                        : (uint64_t)kk * sg.ldb + c)
                       : sg.b_off;
                rb[g][i] = B[idx];
                mask |= (uint32_t)ok << (4 * GA + 4 * g + i);
            }
        }
        okmask = mask;
    }

    // registers -> one LDS buffer: alpha applied to A, masked, split
    __device__ __forceinline__ void store(__bf16 *buf, const Seg &sg,
                                          int tid) const {
        const bool ta = sg.flags & 1, tb = (sg.flags >> 1) & 1;
        const float alpha = (float)sg.alpha;
        __bf16 *Ah = buf, *Al = buf + ASZ, *Bh = buf + 2 * ASZ,
               *Bl = buf + 2 * ASZ + BSZ;
#pragma unroll
        for (int g = 0; g < GA; g++) {
            int m, kg;
            a_coord(ta, tid + g * 256, m, kg);
            float v[4];
#pragma unroll
            for (int i = 0; i < 4; i++)
                v[i] = ((okmask >> (4 * g + i)) & 1)
                           ? (float)ra[g][i] * alpha
                           : 0.0f;
            bf16x4 hi, lo;
            split4(v, hi, lo);
            *(bf16x4 *)(Ah + m * LDK + 4 * kg) = hi;
            *(bf16x4 *)(Al + m * LDK + 4 * kg) = lo;
        }
#pragma unroll
        for (int g = 0; g < GB; g++) {
            int n, kg;
            b_coord(tb, tid + g * 256, n, kg);
            float v[4];
#pragma unroll
            for (int i = 0; i < 4; i++)
                v[i] = ((okmask >> (4 * GA + 4 * g + i)) & 1)
                           ? (float)rb[g][i]
                           : 0.0f;
            bf16x4 hi, lo;
            split4(v, hi, lo);
            *(bf16x4 *)(Bh + n * LDK + 4 * kg) = hi;
            *(bf16x4 *)(Bl + n * LDK + 4 * kg) = lo;
        }
    }
};

template <typename TA, typename TB, typename TD, int TM, int TN>
__global__ __launch_bounds__(256) void seg_gemm_bf3_kernel(
    const TA *__restrict__ A, const TB *__restrict__ B, TD *__restrict__ D,
    TD *__restrict__ P, const Task *__restrict__ tasks, int ntasks,
    const Seg *__restrict__ segs, uint64_t a_rs, uint64_t b_rs, uint64_t d_rs,
    uint64_t p_rs, double scale, int beta0, XcdSplit xs) {
    typedef Bf3Tile<TA, TB, TM, TN> T;
    constexpr int WM = TM / 2, WN = TN / 2;
    constexpr int FM = WM / 16, FN = WN / 16;
    __shared__ __attribute__((aligned(16))) __bf16 smem[2 * T::BUF];
    const int rhs = blockIdx.y;
    A += rhs * a_rs;
    B += rhs * b_rs;
    D += rhs * d_rs;
    P += rhs * p_rs;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int lr = lane & 15, lk = lane >> 4;
    T ld;
    int t0 = blockIdx.x, tend = ntasks, tstep = gridDim.x;
    if (xs.off[8] != 0) { // per-XCD task ranges, as in seg_gemm_kernel
        const int x = blockIdx.x & 7;
        t0 = (int)xs.off[x] + (int)(blockIdx.x >> 3);
        tend = (int)xs.off[x + 1];
        tstep = gridDim.x >> 3;
    }
    for (int t = t0; t < tend; t += tstep) {
        const Task task = tasks[t];
        f32x4 acc[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FN; j++)
                acc[i][j] = f32x4{0, 0, 0, 0};
        uint32_t s = task.seg_begin, k0 = 0;
        if (s < task.seg_end) {
            Seg cur = segs[s];
            Seg nxt = cur;
            if (s + 1 < task.seg_end)
                nxt = segs[s + 1];
            int buf = 0;
            ld.load(A, B, cur, k0, task, tid);
            ld.store(smem, cur, tid);
            __syncthreads();
            while (true) {
                k0 += T::TK;
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
                const __bf16 *Ah = smem + buf * T::BUF;
                const __bf16 *Al = Ah + T::ASZ;
                const __bf16 *Bh = Ah + 2 * T::ASZ;
                const __bf16 *Bl = Bh + T::BSZ;
                bf16x8 ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
                for (int i = 0; i < FM; i++) {
                    const int o = (wm * WM + i * 16 + lr) * T::LDK + 8 * lk;
                    ah[i] = *(const bf16x8 *)(Ah + o);
                    al[i] = *(const bf16x8 *)(Al + o);
                }
#pragma unroll
                for (int j = 0; j < FN; j++) {
                    const int o = (wn * WN + j * 16 + lr) * T::LDK + 8 * lk;
                    bh[j] = *(const bf16x8 *)(Bh + o);
                    bl[j] = *(const bf16x8 *)(Bl + o);
                }
                // small cross terms first, then the leading term
#pragma unroll
                for (int i = 0; i < FM; i++)
#pragma unroll
                    for (int j = 0; j < FN; j++) {
                        acc[i][j] = mfma_bf16(al[i], bh[j], acc[i][j]);
                        acc[i][j] = mfma_bf16(ah[i], bl[j], acc[i][j]);
                        acc[i][j] = mfma_bf16(ah[i], bh[j], acc[i][j]);
                    }
                if (!more)
                    break;
                buf ^= 1;
                ld.store(smem + buf * T::BUF, cur, tid);
                __syncthreads();
            }
            __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FN; j++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const uint32_t row = task.row0 + wm * WM + i * 16 + 4 * lk + r;
                    const uint32_t col = task.col0 + wn * WN + j * 16 + lr;
                    if (row < task.M && col < task.N) {
                        const uint64_t idx =
                            task.d_off + (uint64_t)row * task.ldd + col;
                        const double v = scale * (double)acc[i][j][r];
                        if (task.beta & 2)
                            P[idx] = (TD)v;
                        else if ((task.beta | beta0) & 1)
                            D[idx] = (TD)((double)D[idx] + v);
                        else
                            D[idx] = (TD)v;
                    }
                }
    }
}

template <typename TA, typename TB, typename TD>
hipError_t launch_seg_gemm_bf3(int cfg, const TA *A, const TB *B, TD *D,
                               TD *P, const Task *tasks, int ntasks,
                               const Seg *segs, uint64_t a_rs, uint64_t b_rs,
                               uint64_t d_rs, uint64_t p_rs, double scale,
                               int beta0, int n_rhs, int grid,
                               hipStream_t st, XcdSplit xs) {
    if (ntasks <= 0)
        return hipSuccess;
    dim3 g(task_grid(grid, ntasks, xs), (unsigned)n_rhs);
#define B2G_LAUNCH(TM_, TN_)                                                   \
    hipLaunchKernelGGL((seg_gemm_bf3_kernel<TA, TB, TD, TM_, TN_>), g,         \
                       dim3(256), 0, st, A, B, D, P, tasks, ntasks, segs,      \
                       a_rs, b_rs, d_rs, p_rs, scale, beta0, xs)
    switch (cfg) {
    case 0: B2G_LAUNCH(32, 32); break;
    case 1: B2G_LAUNCH(32, 64); break;
    case 2: B2G_LAUNCH(32, 128); break;
    case 3: B2G_LAUNCH(64, 32); break;
    case 4: B2G_LAUNCH(64, 64); break;
    case 5: B2G_LAUNCH(64, 128); break;
    case 6: B2G_LAUNCH(128, 32); break;
    case 7: B2G_LAUNCH(128, 64); break;
    default: B2G_LAUNCH(128, 128); break;
    }
#undef B2G_LAUNCH
    return hipGetLastError();
}

template hipError_t launch_seg_gemm_bf3<double, float, float>(
    int, const double *, const float *, float *, float *, const Task *, int,
    const Seg *, uint64_t, uint64_t, uint64_t, uint64_t, double, int, int,
    int, hipStream_t, XcdSplit);
template hipError_t launch_seg_gemm_bf3<float, float, double>(
    int, const float *, const float *, double *, double *, const Task *, int,
    const Seg *, uint64_t, uint64_t, uint64_t, uint64_t, double, int, int,
    int, hipStream_t, XcdSplit);

} // namespace b2g
