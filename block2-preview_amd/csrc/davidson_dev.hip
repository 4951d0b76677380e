// Device-resident Davidson iteration (k roots), built to be captured once
// into a hipGraph and replayed: every step reads the subspace size and the
// Ritz data from a DavState in device memory, so one graph serves all
// iterations and the host only reads a small status header per iteration.
//
// Steps follow IterativeMatrixFunctions::davidson
// (src/core/iterative_matrix_functions.hpp:866-1173, DavidsonTypes::Normal):
// sigma of the newest vector; subspace matrix alpha = <b_i|sigma_j>;
// Ritz pairs; rotation of b and sigma; re-check of the roots converged so
// far (the first that fails becomes the current root ck again); residual
// q = sigma_ck - lambda_ck b_ck and the squared-norm test against
// conv + lambda^2 rel^2 (a root converges only once the subspace holds k
// vectors); Olsen (:93-108), Davidson (:66-72) or no preconditioner;
// restart to def_min vectors at def_max; orthogonalisation of q against
// the subspace and append.  A step after which a root converged appends
// nothing: the host then replays the step alone, without a product.  The Ritz problem is solved on the device by cyclic Jacobi in
// round-robin order (m/2 disjoint rotations at a time); the orthogonalisation
// is classical Gram-Schmidt applied twice (as orthogonal as one modified
// pass, with fixed launches).
#include "kernels.hpp"
#include "launch.hpp"

namespace b2g {

namespace {

constexpr int kGrid = 512; // partial-sum blocks of the reductions

__device__ __forceinline__ double bsum(double v, double *sh) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s)
            sh[tid] += sh[tid + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// S[m-1] = sn, when the previous step appended vector m-1
__global__ void put_sigma_kernel(DavState *st, double *__restrict__ S,
                                 const double *__restrict__ sn, size_t n) {
    if (!st->appended)
        return;
    const int m = st->m;
    double *dst = S + (size_t)(m - 1) * n;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = sn[i];
}

// Lower-triangle pairs j <= i < m of the Gram matrix: p = i(i+1)/2 + j;
// per-block partial sums are kPairStride apart.
constexpr int kPairStride = 64 * 65 / 2;

__device__ __forceinline__ void pair_ij(int p, int &i, int &j) {
    i = (int)((sqrt(8.0 * p + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= p)
        i++;
    while (i * (i + 1) / 2 > p)
        i--;
    j = p - i * (i + 1) / 2;
}

// one block per pair: G[i][j] = G[j][i] = sum over blocks (fixed order)
__global__ __launch_bounds__(256) void gram_reduce_kernel(
    DavState *st, const double *__restrict__ partial, int nblk) {
    __shared__ double sh[256];
    const int m = st->m, npair = m * (m + 1) / 2, p = blockIdx.x;
    if (p >= npair)
        return;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256)
        s += partial[(size_t)b * kPairStride + p];
    s = bsum(s, sh);
    if (threadIdx.x == 0) {
        int i, j;
        pair_ij(p, i, j);
        st->G[i * m + j] = s;
        st->G[j * m + i] = s;
    }
}
__global__ __launch_bounds__(256) void ritz_kernel(DavState *st) {
    __shared__ double a[64][65];
    __shared__ double z[64][65];
    __shared__ int pp[32], qq[32];
    __shared__ double cc[32], ss[32];
    __shared__ double red[256];
    __shared__ int order[64];
    const int m = st->m, tid = threadIdx.x;
    const int np = (m + 1) & ~1; // players (a dummy when m is odd)
    for (int e = tid; e < 64 * 64; e += 256) {
        const int i = e / 64, j = e % 64;
        a[i][j] = (i < m && j < m) ? st->G[i * m + j] : 0.0;
        z[i][j] = i == j ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int sweep = 0; sweep < 100 && m > 1; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int e = tid; e < m * m; e += 256) {
            const int i = e / m, j = e % m;
            const double v = a[i][j] * a[i][j];
            tot += v;
            if (i != j)
                off += v;
        }
        off = bsum(off, red);
        tot = bsum(tot, red);
        if (off <= 1E-30 * tot || off == 0.0)
            break;
        for (int r = 0; r < np - 1; r++) {
            if (tid < np / 2) {
                int p, q;
                if (tid == 0)
                    p = np - 1, q = r;
                else
                    p = (r + tid) % (np - 1),
                    q = (r - tid + (np - 1)) % (np - 1);
                if (p > q) {
                    const int t = p;
                    p = q, q = t;
                }
                double c = 1.0, s = 0.0;
                if (q < m) {
                    const double apq = a[p][q];
                    if (fabs(apq) >= 1E-300) {
                        const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
                        const double t = (theta >= 0 ? 1.0 : -1.0) /
                                         (fabs(theta) + sqrt(theta * theta + 1.0));
                        c = 1.0 / sqrt(t * t + 1.0);
                        s = t * c;
                    }
                }
                pp[tid] = p, qq[tid] = q, cc[tid] = c, ss[tid] = s;
            }
            __syncthreads();
            // columns: A <- A J, Z <- Z J
            for (int e = tid; e < (np / 2) * m; e += 256) {
                const int k = e % m, w = e / m;
                const int p = pp[w], q = qq[w];
                if (q < m) {
                    const double c = cc[w], s = ss[w];
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                    const double zkp = z[k][p], zkq = z[k][q];
                    z[k][p] = c * zkp - s * zkq;
                    z[k][q] = s * zkp + c * zkq;
                }
            }
            __syncthreads();
            // rows: A <- J^T A
            for (int e = tid; e < (np / 2) * m; e += 256) {
                const int k = e % m, w = e / m;
                const int p = pp[w], q = qq[w];
                if (q < m) {
                    const double c = cc[w], s = ss[w];
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
            }
            __syncthreads();
        }
    }
    if (tid == 0) {
        // stable ascending order of the diagonal
        for (int i = 0; i < m; i++)
            order[i] = i;
        for (int i = 1; i < m; i++) {
            const int x = order[i];
            int j = i - 1;
            while (j >= 0 && a[order[j]][order[j]] > a[x][x]) {
                order[j + 1] = order[j];
                j--;
            }
            order[j + 1] = x;
        }
    }
    __syncthreads();
    for (int j = tid; j < m; j += 256)
        st->ld[j] = a[order[j]][order[j]];
    for (int e = tid; e < m * m; e += 256) {
        const int j = e / m, i = e % m;
        st->V[j * m + i] = z[i][order[j]];
    }
}

// X_j <- sum_i V[j*m + i] X_i, j < m (in place, rows of n)
typedef double v4d __attribute__((ext_vector_type(4)));

// FP64 MFMA 16x16x4: lane l supplies A[l%16][l/16] and B[l/16][l%16];
// accumulator element r of lane l is D[4r + l/16][l%16].
__device__ __forceinline__ v4d mfma64(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Partial Gram sums <B_i|S_j>, j <= i < m, on FP64 MFMA: wave w of the grid
// takes a contiguous slab of rows and accumulates the 16 x 16 tiles
// (ti, tj), tj <= ti < MT, with K running along the rows (lane l supplies
// B_{16ti + l%16} and S_{16tj + l%16} at row r + l/16).  The four waves of a
// block are summed in LDS; per block: partial[blockIdx * kPairStride + p],
// p = i(i+1)/2 + j.
template <int MT>
__global__ __launch_bounds__(256) void gram_mfma_kernel(
    const DavState *st, const double *__restrict__ X,
    const double *__restrict__ Y, size_t n, double *__restrict__ partial) {
    constexpr int NP = 16 * MT * (16 * MT + 1) / 2;
    __shared__ double red[4][NP];
    const int m = st->m, npair = m * (m + 1) / 2;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    v4d acc[MT][MT];
#pragma unroll
    for (int a = 0; a < MT; a++)
#pragma unroll
        for (int b = 0; b < MT; b++)
            acc[a][b] = v4d{0, 0, 0, 0};
    const size_t nw = (size_t)gridDim.x * 4, gw = (size_t)blockIdx.x * 4 + wave;
    const size_t slab = (((n + nw - 1) / nw) + 3) & ~(size_t)3;
    const size_t rbeg = gw * slab;
    const size_t rend = rbeg + slab < n ? rbeg + slab : n;
#pragma unroll 4
    for (size_t r = rbeg; r < rend; r += 4) {
        const size_t rr = r + lk;
        const bool ok = rr < rend;
        double xa[MT], yb[MT];
#pragma unroll
        for (int t = 0; t < MT; t++) {
            const int i = 16 * t + li;
            xa[t] = (ok && i < m) ? X[(size_t)i * n + rr] : 0.0;
            yb[t] = (ok && i < m) ? Y[(size_t)i * n + rr] : 0.0;
        }
#pragma unroll
        for (int a = 0; a < MT; a++)
#pragma unroll
            for (int b = 0; b <= a; b++)
                if (16 * a < m) // uniform: tiles past the subspace skipped
                    acc[a][b] = mfma64(xa[a], yb[b], acc[a][b]);
    }
    mfma_settle<v4d, MT * MT>(&acc[0][0]);
    for (int p = tid; p < 4 * NP; p += 256)
        (&red[0][0])[p] = 0.0;
    __syncthreads();
#pragma unroll
    for (int a = 0; a < MT; a++)
#pragma unroll
        for (int b = 0; b <= a; b++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = 16 * a + 4 * q + lk, j = 16 * b + li;
                if (i < m && j <= i)
                    red[wave][i * (i + 1) / 2 + j] = acc[a][b][q];
            }
    __syncthreads();
    for (int p = tid; p < npair; p += 256)
        partial[(size_t)blockIdx.x * kPairStride + p] =
            ((red[0][p] + red[1][p]) + red[2][p]) + red[3][p];
}

// X <- V X (X_j = sum_i V[j*m+i] X_i, j < m) for X = S and X = B on FP64
// MFMA: one wave per 16-row chunk, A = V (zero-padded to 16MT x 16MT, in
// LDS), B = the chunk's m x 16 values (registers); every value of the chunk
// is read before any is written, so the rotation is in place.
template <int MT>
__global__ __launch_bounds__(256) void rotate_mfma_kernel(
    const DavState *st, double *__restrict__ S, double *__restrict__ B,
    size_t n) {
    constexpr int MAXM = 16 * MT, KT = 4 * MT;
    __shared__ double vs[MAXM][MAXM + 1];
    const int m = st->m;
    for (int e = threadIdx.x; e < MAXM * MAXM; e += 256) {
        const int j = e / MAXM, i = e % MAXM;
        vs[j][i] = (j < m && i < m) ? st->V[j * m + i] : 0.0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const size_t nchunk = (n + 15) / 16;
    for (size_t c = (size_t)blockIdx.x * 4 + wave; c < nchunk;
         c += (size_t)gridDim.x * 4) {
        const size_t r = c * 16 + li;
        const bool rok = r < n;
#pragma unroll
        for (int pass = 0; pass < 2; pass++) {
            double *X = pass == 0 ? S : B;
            double b[KT];
#pragma unroll
            for (int kk = 0; kk < KT; kk++) {
                const int k = 4 * kk + lk;
                b[kk] = (k < m && rok) ? X[(size_t)k * n + r] : 0.0;
            }
            v4d acc[MT];
#pragma unroll
            for (int t = 0; t < MT; t++) {
                acc[t] = v4d{0, 0, 0, 0};
#pragma unroll
                for (int kk = 0; kk < KT; kk++)
                    if (16 * t < m && 4 * kk < m) // uniform
                        acc[t] = mfma64(vs[16 * t + li][4 * kk + lk], b[kk],
                                        acc[t]);
            }
            mfma_settle<v4d, MT>(&acc[0]);
#pragma unroll
            for (int t = 0; t < MT; t++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int j = 16 * t + 4 * q + lk;
                    if (j < m && rok)
                        X[(size_t)j * n + r] = acc[t][q];
                }
        }
    }
}

// squared residual norms |S_i - ld_i B_i|^2 of the roots i < ck (the ones
// converged so far), per block
__global__ __launch_bounds__(256) void recheck_kernel(
    const DavState *st, const double *__restrict__ B,
    const double *__restrict__ S, size_t n, double *__restrict__ partial) {
    __shared__ double sh[256];
    const int ck = st->ck;
    for (int i = 0; i < ck; i++) {
        const double l = st->ld[i];
        const double *b = B + (size_t)i * n, *sg = S + (size_t)i * n;
        double s = 0.0;
        for (size_t r = blockIdx.x * (size_t)256 + threadIdx.x; r < n;
             r += (size_t)gridDim.x * 256) {
            const double v = sg[r] - l * b[r];
            s += v * v;
        }
        s = bsum(s, sh);
        if (threadIdx.x == 0)
            partial[(size_t)blockIdx.x * 64 + i] = s;
    }
}

// the first root i < ck whose residual fails the test becomes ck again
// (one block; sums in fixed block order)
__global__ __launch_bounds__(256) void recheck_select_kernel(
    DavState *st, const double *__restrict__ partial, int nblk) {
    __shared__ double sh[256];
    const int ck = st->ck;
    int nck = ck;
    for (int i = 0; i < ck; i++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblk; b += 256)
            s += partial[(size_t)b * 64 + i];
        s = bsum(s, sh);
        const double l = st->ld[i];
        if (fabs(s) >= st->conv + l * l * st->rel2) {
            nck = i;
            break;
        }
    }
    if (threadIdx.x == 0)
        st->ck = nck;
}

// q = S_ck - ld_ck B_ck and partial sums of |q|^2
__global__ __launch_bounds__(256) void residual_kernel(
    const DavState *st, const double *__restrict__ B,
    const double *__restrict__ S, double *__restrict__ q, size_t n,
    double *__restrict__ partial) {
    __shared__ double sh[256];
    const int ck = st->ck;
    const double l = st->ld[ck];
    const double *b = B + (size_t)ck * n, *sg = S + (size_t)ck * n;
    double s = 0.0;
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n;
         i += (size_t)gridDim.x * 256) {
        const double v = sg[i] - l * b[i];
        q[i] = v;
        s += v * v;
    }
    s = bsum(s, sh);
    if (threadIdx.x == 0)
        partial[blockIdx.x] = s;
}

// out[i] = sum_b partial[b*k + i], i < k (one block)
__global__ __launch_bounds__(256) void reduce_kernel(
    const double *__restrict__ partial, int nblk, int k,
    double *__restrict__ out) {
    __shared__ double sh[256];
    for (int i = 0; i < k; i++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblk; b += 256)
            s += partial[(size_t)b * k + i];
        s = bsum(s, sh);
        if (threadIdx.x == 0)
            out[i] = s;
    }
}

// Preconditioner, first pass (root ck): Olsen's t = c / (ld - aa),
// q = q / (ld - aa) where |ld - aa| > 1e-12, with partial sums of c.q and
// c.t; Davidson's (type bit 32) only divides q; NoPrecond (64) leaves q
__global__ __launch_bounds__(256) void olsen1_kernel(
    const DavState *st, double *__restrict__ q, const double *__restrict__ B,
    const double *__restrict__ aa, double *__restrict__ t, size_t n,
    double *__restrict__ partial) {
    __shared__ double sh[256];
    const int ck = st->ck, type = st->type;
    if (type & 64)
        return;
    const double ld = st->ld[ck];
    const double *c = B + (size_t)ck * n;
    double s_cq = 0.0, s_ct = 0.0;
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n;
         i += (size_t)gridDim.x * 256) {
        const double ci = c[i];
        double qi = q[i], ti = ci;
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
    if (type & 32)
        return;
    s_cq = bsum(s_cq, sh);
    s_ct = bsum(s_ct, sh);
    if (threadIdx.x == 0) {
        partial[(size_t)blockIdx.x * 2 + 0] = s_cq;
        partial[(size_t)blockIdx.x * 2 + 1] = s_ct;
    }
}

__global__ void olsen2_kernel(const DavState *st, double *__restrict__ q,
                              const double *__restrict__ t, size_t n) {
    if (st->type & (32 | 64))
        return;
    const double a = -st->scal[1] / st->scal[2];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        q[i] += a * t[i];
}

// convergence test of root ck and restart (one thread)
__global__ void control_kernel(DavState *st) {
    st->iter++;
    const int ck = st->ck;
    st->qq = st->scal[0];
    st->energy = st->ld[0];
    st->appended = 0;
    const double l = st->ld[ck];
    if (fabs(st->qq) < st->conv + l * l * st->rel2 && st->m >= st->k) {
        st->ck = ck + 1;
        if (ck + 1 == st->k)
            st->status = 1;
    } else {
        if (st->m >= st->def_max)
            st->m = st->def_min;
        st->appended = 1;
    }
    if (st->status == 0 && st->iter == st->soft_max_iter)
        st->status = 3;
    else if (st->status == 0 && st->iter >= st->max_iter)
        st->status = 2;
}

// partial dots <B_j|q>, j < m
__global__ __launch_bounds__(256) void gs_dots_kernel(
    const DavState *st, const double *__restrict__ B,
    const double *__restrict__ q, size_t n, double *__restrict__ partial) {
    __shared__ double sh[256];
    if (!st->appended)
        return;
    const int m = st->m;
    for (int j = 0; j < m; j++) {
        double s = 0.0;
        for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n;
             i += (size_t)gridDim.x * 256)
            s += B[(size_t)j * n + i] * q[i];
        s = bsum(s, sh);
        if (threadIdx.x == 0)
            partial[(size_t)blockIdx.x * 64 + j] = s;
    }
}

__global__ __launch_bounds__(256) void gs_reduce_kernel(
    DavState *st, const double *__restrict__ partial, int nblk) {
    __shared__ double sh[256];
    if (!st->appended)
        return;
    const int m = st->m;
    for (int j = 0; j < m; j++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblk; b += 256)
            s += partial[(size_t)b * 64 + j];
        s = bsum(s, sh);
        if (threadIdx.x == 0)
            st->dots[j] = s;
    }
}

// q -= sum_j dots_j B_j
__global__ __launch_bounds__(256) void gs_update_kernel(
    const DavState *st, const double *__restrict__ B, double *__restrict__ q,
    size_t n) {
    __shared__ double d[64];
    if (!st->appended)
        return;
    const int m = st->m;
    for (int j = threadIdx.x; j < m; j += 256)
        d[j] = st->dots[j];
    __syncthreads();
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n;
         i += (size_t)gridDim.x * 256) {
        double s = q[i];
        for (int j = 0; j < m; j++)
            s -= d[j] * B[(size_t)j * n + i];
        q[i] = s;
    }
}

__global__ __launch_bounds__(256) void norm_kernel(
    const DavState *st, const double *__restrict__ q, size_t n,
    double *__restrict__ partial) {
    __shared__ double sh[256];
    if (!st->appended)
        return;
    double s = 0.0;
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n;
         i += (size_t)gridDim.x * 256)
        s += q[i] * q[i];
    s = bsum(s, sh);
    if (threadIdx.x == 0)
        partial[blockIdx.x] = s;
}

// q /= |q|; qn = q; B[m] = q
__global__ void append_kernel(const DavState *st, const double *__restrict__ q,
                              double *__restrict__ qn, double *__restrict__ B,
                              size_t n) {
    if (!st->appended)
        return;
    const double f = 1.0 / sqrt(st->scal[3]);
    double *dst = B + (size_t)st->m * n;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const double v = q[i] * f;
        qn[i] = v;
        dst[i] = v;
    }
}

// the last kernel of a step: grow the subspace, then publish the 128-byte
// state header to the host (pinned, device-mapped memory)
__global__ void bump_kernel(DavState *st, DavState *host) {
    if (st->appended)
        st->m++;
    if (host != nullptr) {
        const int *src = (const int *)st;
        volatile int *dst = (volatile int *)host;
        for (int k = 0; k < 32; k++)
            dst[k] = src[k];
        __threadfence_system();
    }
}

dim3 vgrid(size_t n) {
    size_t g = (n + 255) / 256;
    if (g > 8192)
        g = 8192;
    return dim3((unsigned)(g ? g : 1));
}

} // namespace

// The Davidson step after the matvec of the newest vector (sn = H qn):
// everything up to and including the append of the next vector.  All
// launches have fixed arguments; sizes come from the state.
hipError_t launch_davidson_step(DavState *st, double *B, double *S,
                                const double *sn, double *q, double *t,
                                double *qn, const double *diag, size_t n,
                                int nv, double *partial, DavState *host,
                                hipStream_t s) {
    if (nv < 2 || nv > 64)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(put_sigma_kernel, vgrid(n), dim3(256), 0, s, st, S, sn,
                       n);
    const int mt = (nv + 15) / 16;
#define B2G_MT(K, ...)                                                         \
    switch (mt) {                                                              \
    case 1: hipLaunchKernelGGL(K<1>, __VA_ARGS__); break;                      \
    case 2: hipLaunchKernelGGL(K<2>, __VA_ARGS__); break;                      \
    case 3: hipLaunchKernelGGL(K<3>, __VA_ARGS__); break;                      \
    default: hipLaunchKernelGGL(K<4>, __VA_ARGS__); break;                     \
    }
    B2G_MT(gram_mfma_kernel, dim3(kGrid), dim3(256), 0, s, st, B, S, n,
           partial);
    hipLaunchKernelGGL(gram_reduce_kernel, dim3(kPairStride), dim3(256), 0, s,
                       st, partial, kGrid);
    hipLaunchKernelGGL(ritz_kernel, dim3(1), dim3(256), 0, s, st);
    B2G_MT(rotate_mfma_kernel, vgrid(n), dim3(256), 0, s, st, S, B, n);
#undef B2G_MT
    hipLaunchKernelGGL(recheck_kernel, dim3(kGrid), dim3(256), 0, s, st, B, S,
                       n, partial);
    hipLaunchKernelGGL(recheck_select_kernel, dim3(1), dim3(256), 0, s, st,
                       partial, kGrid);
    hipLaunchKernelGGL(residual_kernel, dim3(kGrid), dim3(256), 0, s, st, B,
                       S, q, n, partial);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(256), 0, s, partial, kGrid,
                       1, st->scal + 0);
    hipLaunchKernelGGL(olsen1_kernel, dim3(kGrid), dim3(256), 0, s, st, q, B,
                       diag, t, n, partial);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(256), 0, s, partial, kGrid,
                       2, st->scal + 1);
    hipLaunchKernelGGL(olsen2_kernel, vgrid(n), dim3(256), 0, s, st, q, t, n);
    hipLaunchKernelGGL(control_kernel, dim3(1), dim3(1), 0, s, st);
    for (int pass = 0; pass < 2; pass++) {
        hipLaunchKernelGGL(gs_dots_kernel, dim3(kGrid), dim3(256), 0, s, st, B,
                           q, n, partial);
        hipLaunchKernelGGL(gs_reduce_kernel, dim3(1), dim3(256), 0, s, st,
                           partial, kGrid);
        hipLaunchKernelGGL(gs_update_kernel, vgrid(n), dim3(256), 0, s, st, B,
                           q, n);
    }
    hipLaunchKernelGGL(norm_kernel, dim3(kGrid), dim3(256), 0, s, st, q, n,
                       partial);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(256), 0, s, partial, kGrid,
                       1, st->scal + 3);
    hipLaunchKernelGGL(append_kernel, vgrid(n), dim3(256), 0, s, st, q, qn, B,
                       n);
    hipLaunchKernelGGL(bump_kernel, dim3(1), dim3(1), 0, s, st, host);
    return hipGetLastError();
}
int davidson_partial_len() { return kGrid * 64 * 64; }

} // namespace b2g
