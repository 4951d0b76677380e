// Segmented grouped GEMM in split-BF16 three-term arithmetic (gfx950) on
// pre-split operands.
//
// Every operand value a is split once into two BF16 numbers, hi = bf16(a)
// and lo = bf16(a - hi), so that a = hi + lo to about 2^-17 relative.  A
// product is then
//     a * b ~= a_hi * b_hi + a_hi * b_lo + a_lo * b_hi
// (the dropped a_lo * b_lo is ~2^-16 relative), each term one
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  Three BF16 MFMAs cost
// 3 * 16 cycles per 16x16x32 block against 8 * 32 cycles for the same block
// with FP32 16x16x4 MFMAs; the Davidson that uses them refines in FP64.
//
// Packed planes.  A GEMM operand is held as a row-major R x ld matrix whose
// rows run along the contraction index k (ld = K rounded up to 8, columns
// [K, ld) zero), stored twice with identical layout: the hi plane and the lo
// plane.  Eight consecutive k of a row are one aligned 16-byte vector, so the
// tile loader issues 16-byte buffer loads and 16-byte LDS stores and does no
// arithmetic.  The constant operators are packed once per plan (in the
// orientation each pair reads them), the wavefunction once per H*psi, and
// phase 1 writes its product W directly as packed W^T planes for phase 2.
//
// Task / segment semantics are those of seg_gemm_kernel (kernels.hip): a
// task is one TM x TN output tile accumulating sum_s alpha_s A_s B_s^T over
// its segments (A_s: M x K_s, B_s: N x K_s, both packed), the tile owned by
// exactly one task (no atomics).  alpha_s is applied to each segment's FP32
// partial sum when the segment ends.
#include "kernels.hpp"
#include "launch.hpp"

namespace b2g {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Lane l of v_mfma_f32_16x16x32_bf16 supplies A[l%16][8*(l/16) + i] and
// B[8*(l/16) + i][l%16] (i < 8); accumulator element r of lane l is
// D[4*(l/16) + r][l%16] (in-bounds probe of the builtin on gfx950).
__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// raw buffer resource over one packed matrix: offsets are 32-bit byte
// offsets below 2^31 (matrices hold fewer than 2^28 elements); an offset of
// kPkMasked lies past the record count and reads as zero
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pk_rsrc(const __bf16 *p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<__bf16 *>(p),
                                             (short)0, 0x7FFFFFF0, 0x00020000);
}
constexpr uint32_t kPkMasked = 0x80000000u;

__device__ __forceinline__ u32x4 pk_load(__amdgpu_buffer_rsrc_t r,
                                         uint32_t off) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
}

// ------------------------------------------------------------------ packing
// One thread per 8-element group (8 consecutive k of one packed row).  The
// groups of a job are numbered matrix after matrix from each matrix's g0:
// row-major (row, k group) when the source runs along k, column-major
// (k group, row) for a transposed source, so that consecutive lanes read
// consecutive source addresses.  A workgroup takes 256 * kPackGPT
// consecutive groups; the group offsets of the matrices they span (at most
// one matrix per group) are staged in LDS for each thread's binary search.
constexpr int kPackGPT = 4;

__global__ __launch_bounds__(256) void pack_bf16_kernel(
    const double *__restrict__ src, __bf16 *__restrict__ hi,
    __bf16 *__restrict__ lo, const PackMat *__restrict__ mats, int nmats,
    uint64_t ngroups, uint64_t src_rs, uint64_t dst_rs) {
    constexpr int G = 256 * kPackGPT;
    __shared__ uint64_t sg0[G + 1];
    __shared__ int sm[2];
    const int rhs = blockIdx.y;
    src += rhs * src_rs;
    hi += rhs * dst_rs;
    lo += rhs * dst_rs;
    const uint64_t base = (uint64_t)blockIdx.x * G;
    if (base >= ngroups)
        return;
    const uint64_t last = (base + G < ngroups ? base + G : ngroups) - 1;
    if (threadIdx.x < 2) {
        // last matrix with g0 <= base (thread 0) / <= last (thread 1)
        const uint64_t key = threadIdx.x == 0 ? base : last;
        int a = 0, b = nmats - 1;
        while (a < b) {
            const int mid = (a + b + 1) >> 1;
            if (mats[mid].g0 <= key)
                a = mid;
            else
                b = mid - 1;
        }
        sm[threadIdx.x] = a;
    }
    __syncthreads();
    const int m0 = sm[0], nm = sm[1] - sm[0] + 1; // nm <= G
    for (int i = threadIdx.x; i < nm; i += 256)
        sg0[i] = mats[m0 + i].g0;
    __syncthreads();
    for (int it = 0; it < kPackGPT; it++) {
        const uint64_t g = base + (uint64_t)it * 256 + threadIdx.x;
        if (g > last)
            break;
        int a = 0, b = nm - 1;
        while (a < b) {
            const int mid = (a + b + 1) >> 1;
            if (sg0[mid] <= g)
                a = mid;
            else
                b = mid - 1;
        }
        const PackMat pm = mats[m0 + a];
        const uint32_t l = (uint32_t)(g - pm.g0);
        const uint32_t kg = pm.dst_ld >> 3;
        uint32_t r, k;
        if (pm.trans)
            k = (l / pm.rows) * 8, r = l % pm.rows;
        else
            r = l / kg, k = (l % kg) * 8;
        bf16x8 h, lw;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint32_t kk = k + i;
            double v = 0.0;
            if (kk < pm.cols)
                v = src[pm.src + (pm.trans ? (uint64_t)kk * pm.src_ld + r
                                           : (uint64_t)r * pm.src_ld + kk)];
            const __bf16 vh = (__bf16)(float)v;
            h[i] = vh;
            lw[i] = (__bf16)(float)(v - (double)(float)vh);
        }
        const uint64_t o = pm.dst + (uint64_t)r * pm.dst_ld + k;
        *(bf16x8 *)(hi + o) = h;
        *(bf16x8 *)(lo + o) = lw;
    }
}

hipError_t launch_pack_bf16(const double *src, __bf16 *hi, __bf16 *lo,
                            const PackMat *mats, int nmats, uint64_t ngroups,
                            uint64_t src_rs, uint64_t dst_rs, int n_rhs,
                            int grid, hipStream_t st) {
    (void)grid;
    if (nmats <= 0 || ngroups == 0)
        return hipSuccess;
    const uint64_t g = (ngroups + 256 * kPackGPT - 1) / (256 * kPackGPT);
    if (g >= (1ull << 31))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(pack_bf16_kernel, dim3((unsigned)g, (unsigned)n_rhs),
                       dim3(256), 0, st, src, hi, lo, mats, nmats, ngroups,
                       src_rs, dst_rs);
    return hipGetLastError();
}

// ------------------------------------------------------------- tile loader
// One workgroup = 4 waves (2x2); K staged through two LDS buffers in steps
// of 32 (one MFMA k-step).  An LDS row holds the 32 k values of one tile row
// (A) or tile column (B) plus 8 of padding, so that a lane's 8 k values are
// one 16-byte read and the 16 rows of a fragment fall on distinct banks.
template <int TM, int TN> struct PkTile {
    static constexpr int TK = 32, LDK = TK + 8;
    static constexpr int VA = TM * TK / 8, VB = TN * TK / 8; // 16 B vectors
    static constexpr int GA = (VA + 255) / 256, GB = (VB + 255) / 256;
    static constexpr int ASZ = TM * LDK, BSZ = TN * LDK;
    static constexpr int BUF = 2 * ASZ + 2 * BSZ; // A hi, A lo, B hi, B lo
    u32x4 ah[GA], al[GA], bh[GB], bl[GB];

    // vector e of a tile: row and k group.  Sixteen consecutive lanes take
    // one k group of sixteen consecutive rows, so that a wave's 16-byte LDS
    // stores follow the bank pattern of the fragment reads (rows 80 bytes
    // apart: any 16 of them cover distinct 16-byte bank groups of 256
    // bytes); a wave still covers 16 whole rows of 64 bytes per global load.
    static __device__ __forceinline__ int vrow(int e) {
        return (e & 15) + 16 * (e >> 6);
    }
    static __device__ __forceinline__ int vgrp(int e) { return (e >> 4) & 3; }

    __device__ __forceinline__ void load(
        __amdgpu_buffer_rsrc_t rAh, __amdgpu_buffer_rsrc_t rAl,
        __amdgpu_buffer_rsrc_t rBh, __amdgpu_buffer_rsrc_t rBl, const Seg &sg,
        uint32_t k0, const Task &task, int tid) {
#pragma unroll
        for (int g = 0; g < GA; g++) {
            const int e = tid + 256 * g;
            if (VA % 256 == 0 || e < VA) {
                const uint32_t r = task.row0 + vrow(e);
                const uint32_t k = k0 + 8 * vgrp(e);
                const uint32_t off = (r < task.M && k < sg.K)
                                         ? 2 * (r * sg.lda + k)
                                         : kPkMasked;
                ah[g] = pk_load(rAh, off);
                al[g] = pk_load(rAl, off);
            }
        }
#pragma unroll
        for (int g = 0; g < GB; g++) {
            const int e = tid + 256 * g;
            if (VB % 256 == 0 || e < VB) {
                const uint32_t c = task.col0 + vrow(e);
                const uint32_t k = k0 + 8 * vgrp(e);
                const uint32_t off = (c < task.N && k < sg.K)
                                         ? 2 * (c * sg.ldb + k)
                                         : kPkMasked;
                bh[g] = pk_load(rBh, off);
                bl[g] = pk_load(rBl, off);
            }
        }
    }

    __device__ __forceinline__ void store(__bf16 *buf, int tid) const {
#pragma unroll
        for (int g = 0; g < GA; g++) {
            const int e = tid + 256 * g;
            if (VA % 256 == 0 || e < VA) {
                const int o = vrow(e) * LDK + 8 * vgrp(e);
                *(u32x4 *)(buf + o) = ah[g];
                *(u32x4 *)(buf + ASZ + o) = al[g];
            }
        }
#pragma unroll
        for (int g = 0; g < GB; g++) {
            const int e = tid + 256 * g;
            if (VB % 256 == 0 || e < VB) {
                const int o = vrow(e) * LDK + 8 * vgrp(e);
                *(u32x4 *)(buf + 2 * ASZ + o) = bh[g];
                *(u32x4 *)(buf + 2 * ASZ + BSZ + o) = bl[g];
            }
        }
    }
};

// OUT 0: D0/D1 are the hi/lo planes of the packed transpose of the output
//        (element (m, n) of a task's block at d_off + n*ldd + m, ldd the
//        packed row length; rows M <= m < ldd receive the zeros of the
//        masked A rows, so the padding of W^T is rewritten by every launch),
//        written through LDS as 16-byte vectors along m;
// OUT 1: D0 is the FP64 output (row-major, ldd), D1 the FP64 split-K
//        partial buffer (task.beta bit 1), as in seg_gemm_kernel.
template <int TM, int TN, int OUT>
__global__ __launch_bounds__(256) void seg_gemm_pk_kernel(
    const __bf16 *__restrict__ Ah, const __bf16 *__restrict__ Al,
    const __bf16 *__restrict__ Bh, const __bf16 *__restrict__ Bl,
    void *__restrict__ D0, void *__restrict__ D1,
    const Task *__restrict__ tasks, int ntasks, const Seg *__restrict__ segs,
    uint64_t a_rs, uint64_t b_rs, uint64_t d_rs, uint64_t p_rs, double scale,
    int beta0, XcdSplit xs) {
    typedef PkTile<TM, TN> T;
    constexpr int WM = TM / 2, WN = TN / 2;
    constexpr int FM = WM / 16, FN = WN / 16;
    __shared__ __attribute__((aligned(16))) __bf16 smem[2 * T::BUF];
    const int rhs = blockIdx.y;
    Ah += rhs * a_rs;
    Al += rhs * a_rs;
    Bh += rhs * b_rs;
    Bl += rhs * b_rs;
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
        // acc: the tile's sum; part: the running segment's unscaled sum
        // (phase 1 tasks have one segment with alpha = 1: part is the sum)
        f32x4 acc[FM][FN], part[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FN; j++) {
                acc[i][j] = f32x4{0, 0, 0, 0};
                part[i][j] = f32x4{0, 0, 0, 0};
            }
        uint32_t s = task.seg_begin, k0 = 0;
        if (s < task.seg_end) {
            Seg cur = segs[s];
            Seg nxt = cur;
            if (s + 1 < task.seg_end)
                nxt = segs[s + 1];
            auto rs = [&](const Seg &sg, __amdgpu_buffer_rsrc_t &a0,
                          __amdgpu_buffer_rsrc_t &a1,
                          __amdgpu_buffer_rsrc_t &b0,
                          __amdgpu_buffer_rsrc_t &b1) {
                a0 = pk_rsrc(Ah + sg.a_off);
                a1 = pk_rsrc(Al + sg.a_off);
                b0 = pk_rsrc(Bh + sg.b_off);
                b1 = pk_rsrc(Bl + sg.b_off);
            };
            __amdgpu_buffer_rsrc_t rAh, rAl, rBh, rBl;
            rs(cur, rAh, rAl, rBh, rBl);
            int buf = 0;
            ld.load(rAh, rAl, rBh, rBl, cur, k0, task, tid);
            ld.store(smem, tid);
            __syncthreads();
            float alpha = (float)cur.alpha; // of the segment in LDS
            while (true) {
                k0 += T::TK;
                const bool seg_done = k0 >= cur.K;
                const float a_done = alpha;
                if (seg_done) {
                    k0 = 0;
                    s++;
                    cur = nxt;
                    if (s + 1 < task.seg_end)
                        nxt = segs[s + 1];
                    rs(cur, rAh, rAl, rBh, rBl);
                    alpha = (float)cur.alpha;
                }
                const bool more = s < task.seg_end;
                if (more)
                    ld.load(rAh, rAl, rBh, rBl, cur, k0, task, tid);
                const __bf16 *sAh = smem + buf * T::BUF;
                const __bf16 *sAl = sAh + T::ASZ;
                const __bf16 *sBh = sAh + 2 * T::ASZ;
                const __bf16 *sBl = sBh + T::BSZ;
                bf16x8 fah[FM], fal[FM], fbh[FN], fbl[FN];
#pragma unroll
                for (int i = 0; i < FM; i++) {
                    const int o = (wm * WM + i * 16 + lr) * T::LDK + 8 * lk;
                    fah[i] = *(const bf16x8 *)(sAh + o);
                    fal[i] = *(const bf16x8 *)(sAl + o);
                }
#pragma unroll
                for (int j = 0; j < FN; j++) {
                    const int o = (wn * WN + j * 16 + lr) * T::LDK + 8 * lk;
                    fbh[j] = *(const bf16x8 *)(sBh + o);
                    fbl[j] = *(const bf16x8 *)(sBl + o);
                }
                // small cross terms first, then the leading term
#pragma unroll
                for (int i = 0; i < FM; i++)
#pragma unroll
                    for (int j = 0; j < FN; j++) {
                        part[i][j] = mfma_bf16(fal[i], fbh[j], part[i][j]);
                        part[i][j] = mfma_bf16(fah[i], fbl[j], part[i][j]);
                        part[i][j] = mfma_bf16(fah[i], fbh[j], part[i][j]);
                    }
                if (OUT == 1 && seg_done) {
                    mfma_settle<f32x4, FM * FN>(&part[0][0]);
#pragma unroll
                    for (int i = 0; i < FM; i++)
#pragma unroll
                        for (int j = 0; j < FN; j++) {
                            acc[i][j] += a_done * part[i][j];
                            part[i][j] = f32x4{0, 0, 0, 0};
                        }
                }
                if (!more)
                    break;
                buf ^= 1;
                ld.store(smem + buf * T::BUF, tid);
                __syncthreads();
            }
            __syncthreads();
        }
        if constexpr (OUT == 0) {
            // W^T through LDS: the tile (row m, column n) lands in a float
            // [TM][TN + 4] array in the (free) staging buffers, then every
            // thread writes 8 consecutive m of one W^T row as a 16-byte
            // vector per plane (whole rows of 256 bytes per 16 lanes)
            mfma_settle<f32x4, FM * FN>(&part[0][0]);
            constexpr int LDT = TN + 4;
            static_assert(TM * LDT * 4 <= 2 * T::BUF * 2, "W^T staging");
            float *tile = (float *)smem;
            __syncthreads();
#pragma unroll
            for (int i = 0; i < FM; i++)
#pragma unroll
                for (int j = 0; j < FN; j++)
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        tile[(wm * WM + i * 16 + 4 * lk + r) * LDT + wn * WN +
                             j * 16 + lr] = part[i][j][r];
            __syncthreads();
            __bf16 *Wh = (__bf16 *)D0 + rhs * d_rs;
            __bf16 *Wl = (__bf16 *)D1 + rhs * d_rs;
            constexpr int MG = TM / 8;
            for (int v = tid; v < TN * MG; v += 256) {
                const int n = v / MG, m0 = 8 * (v % MG);
                const uint32_t row = task.row0 + m0, col = task.col0 + n;
                if (row < task.ldd && col < task.N) {
                    bf16x8 h, l;
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        const float x = tile[(m0 + q) * LDT + n];
                        const __bf16 xh = (__bf16)x;
                        h[q] = xh;
                        l[q] = (__bf16)(x - (float)xh);
                    }
                    const uint64_t o =
                        task.d_off + (uint64_t)col * task.ldd + row;
                    *(bf16x8 *)(Wh + o) = h;
                    *(bf16x8 *)(Wl + o) = l;
                }
            }
            __syncthreads(); // the staging buffers are the next task's
        } else {
            double *D = (double *)D0 + rhs * d_rs;
            double *P = (double *)D1 + rhs * p_rs;
#pragma unroll
            for (int i = 0; i < FM; i++)
#pragma unroll
                for (int j = 0; j < FN; j++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const uint32_t row =
                            task.row0 + wm * WM + i * 16 + 4 * lk + r;
                        const uint32_t col = task.col0 + wn * WN + j * 16 + lr;
                        if (row < task.M && col < task.N) {
                            const uint64_t idx =
                                task.d_off + (uint64_t)row * task.ldd + col;
                            const double v = scale * (double)acc[i][j][r];
                            if (task.beta & 2)
                                P[idx] = v;
                            else if ((task.beta | beta0) & 1)
                                D[idx] += v;
                            else
                                D[idx] = v;
                        }
                    }
        }
    }
}

template <int OUT>
hipError_t launch_seg_gemm_pk(int cfg, const __bf16 *Ah, const __bf16 *Al,
                              const __bf16 *Bh, const __bf16 *Bl, void *D0,
                              void *D1, const Task *tasks, int ntasks,
                              const Seg *segs, uint64_t a_rs, uint64_t b_rs,
                              uint64_t d_rs, uint64_t p_rs, double scale,
                              int beta0, int n_rhs, int grid, hipStream_t st,
                              XcdSplit xs) {
    if (ntasks <= 0)
        return hipSuccess;
    dim3 g(task_grid(grid, ntasks, xs), (unsigned)n_rhs);
#define B2G_LAUNCH(TM_, TN_)                                                   \
    hipLaunchKernelGGL((seg_gemm_pk_kernel<TM_, TN_, OUT>), g, dim3(256), 0,   \
                       st, Ah, Al, Bh, Bl, D0, D1, tasks, ntasks, segs, a_rs,  \
                       b_rs, d_rs, p_rs, scale, beta0, xs)
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

template hipError_t launch_seg_gemm_pk<0>(int, const __bf16 *, const __bf16 *,
                                          const __bf16 *, const __bf16 *,
                                          void *, void *, const Task *, int,
                                          const Seg *, uint64_t, uint64_t,
                                          uint64_t, uint64_t, double, int, int,
                                          int, hipStream_t, XcdSplit);
template hipError_t launch_seg_gemm_pk<1>(int, const __bf16 *, const __bf16 *,
                                          const __bf16 *, const __bf16 *,
                                          void *, void *, const Task *, int,
                                          const Seg *, uint64_t, uint64_t,
                                          uint64_t, uint64_t, double, int, int,
                                          int, hipStream_t, XcdSplit);

} // namespace b2g
