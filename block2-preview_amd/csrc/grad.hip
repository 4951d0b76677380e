// Partial expectations: gradients of E = <v| H |c> with respect to every
// element of the left and right operator buffers, on the grouped GEMM.
//
// For one recorded pair (V[iv] += alpha op(A) C op(B)^T, the rotate of
// src/core/operator_functions.hpp:509-512) the contribution to E is
//     alpha * tr(V^T op(A) C op(B)^T),      V = the bra block of v,
// which is linear in A and in B.  Its derivatives, in each block's stored
// layout (the quantities tensor_left/right_partial_expectation contract,
// src/core/operator_functions.hpp:354-472):
//   dE/dA = alpha V W^T          (A stored v_m x c_m)   W = C op(B)^T
//         = alpha W V^T          (A stored c_m x v_m, conj bit 0)
//   dE/dB = alpha Z^T C          (B stored v_n x c_n)   Z = op(A)^T V
//         = alpha C^T Z          (B stored c_n x v_n, conj bit 1)
// C and V may be windows with row strides c_ld / v_ld (three-tensor
// pairs); their factors are constants of the plan.
// Stage 1 computes every distinct W and Z once into scratch; stage 2 sums,
// per operator block, all its pairs' products along K (output-stationary,
// no atomics, deterministic), in four launches split by transpose and by
// which buffer the left operand comes from.  Gradients sum over all pairs
// that reference a block, so a block shared by several strings gets the
// total derivative.
#include "replay.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <tuple>

namespace b2g {

struct Launch {
    std::vector<Task> t[kTileConfigs];
    std::vector<Seg> s;
    Task *d_t[kTileConfigs] = {};
    Seg *d_s = nullptr;
};

struct GradChunk {
    Launch w, z;           // stage 1
    Launch l0, l1, r0, r1; // stage 2 (see launch order in b2g_grad_ops)
    uint64_t w_len = 0, z_len = 0;
};

struct GradPlan {
    std::vector<GradChunk> chunks;
    uint64_t s_max = 0; // scratch elements (W + Z) of the largest chunk
    void *dev = nullptr;
    double *scratch = nullptr;
    double flops = 0; // algorithmic FLOPs per right-hand side
};

void grad_plan_destroy(GradPlan *g) {
    if (g == nullptr)
        return;
    if (g->dev)
        hipFree(g->dev);
    if (g->scratch)
        hipFree(g->scratch);
    delete g;
}

static Seg make_seg(uint64_t a_off, uint32_t lda, bool ta, uint64_t b_off,
                    uint32_t ldb, bool tb, uint32_t K, double alpha) {
    Seg s;
    s.a_off = a_off;
    s.lda = lda;
    s.b_off = b_off;
    s.ldb = ldb;
    s.K = K;
    s.flags = (ta ? 1u : 0u) | (tb ? 2u : 0u);
    s.alpha = alpha;
    return s;
}

// stage-2 pass: one output block per key, its pairs' segments concatenated
template <typename SegOf>
static void add_pass(Launch &L, const std::map<uint64_t, std::vector<size_t>> &by,
                     const b2g_pair *pairs, SegOf seg_of) {
    for (const auto &kv : by) {
        const uint32_t sb = (uint32_t)L.s.size();
        uint64_t K = 0;
        uint32_t rows = 0, cols = 0;
        for (size_t i : kv.second) {
            const Seg s = seg_of(pairs[i], i, rows, cols);
            K += s.K;
            L.s.push_back(s);
        }
        add_tiles(L.t, kv.first, rows, cols, K, sb, (uint32_t)L.s.size(), 1);
    }
}

// Every element a launch can touch lies inside its buffers: for each task,
// each segment's last A and B element (given the task block's rows M and
// columns N) and the block's last output element.
static bool launch_in_bounds(const Launch &L, uint64_t a_len, uint64_t b_len,
                             uint64_t d_len) {
    for (int c = 0; c < kTileConfigs; c++)
        for (const Task &t : L.t[c]) {
            if (t.M == 0 || t.N == 0)
                continue;
            if (t.d_off + (uint64_t)(t.M - 1) * t.ldd + (t.N - 1) >= d_len)
                return false;
            for (uint32_t k = t.seg_begin; k < t.seg_end; k++) {
                const Seg &s = L.s[k];
                if (s.K == 0)
                    continue;
                const bool ta = s.flags & 1, tb = s.flags & 2;
                const uint64_t amax =
                    s.a_off + (ta ? (uint64_t)(s.K - 1) * s.lda + (t.M - 1)
                                  : (uint64_t)(t.M - 1) * s.lda + (s.K - 1));
                const uint64_t bmax =
                    s.b_off + (tb ? (uint64_t)(t.N - 1) * s.ldb + (s.K - 1)
                                  : (uint64_t)(s.K - 1) * s.ldb + (t.N - 1));
                // 32-bit byte offsets within a segment's blocks
                const uint64_t lim = ((uint64_t)1 << 31) / 8 - 2;
                if (amax >= a_len || bmax >= b_len || amax - s.a_off >= lim ||
                    bmax - s.b_off >= lim)
                    return false;
            }
        }
    return true;
}

static int grad_plan_build(b2g_plan p, GradPlan **out) {
    tile_env_refresh();
    GradPlan *g = new GradPlan();
    const b2g_pair *pairs = p->pairs.data();
    const size_t n = p->pairs.size();
    const uint64_t cap = w_cap_elems();
    size_t i0 = 0;
    while (i0 < n) {
        GradChunk ch;
        // (operator block, vector window start, transpose bit, window rows,
        // window columns, row stride)
        typedef std::tuple<uint64_t, uint64_t, uint32_t, uint32_t, uint32_t,
                           uint32_t>
            Key;
        std::map<Key, uint64_t> wmap, zmap;
        std::map<size_t, std::pair<uint64_t, uint64_t>> pwz; // pair -> (W, Z)
        size_t i1 = i0;
        for (; i1 < n; i1++) {
            const b2g_pair &q = pairs[i1];
            if (q.c_m == 0 || q.c_n == 0 || q.v_m == 0 || q.v_n == 0)
                continue;
            const Key kw(q.b_off, q.c_off, q.conj & 2, q.c_m, q.c_n, q.c_ld);
            const Key kz(q.a_off, q.v_off, q.conj & 1, q.v_m, q.v_n, q.v_ld);
            const uint64_t sz = (uint64_t)q.c_m * q.v_n;
            const uint64_t need = (wmap.count(kw) ? 0 : sz) +
                                  (zmap.count(kz) ? 0 : sz);
            if (ch.w_len + ch.z_len + need > cap && i1 > i0)
                break;
            // row strides of windowed C / V (three-tensor pairs)
            const uint32_t cld = q.c_ld ? q.c_ld : q.c_n;
            const uint32_t vld = q.v_ld ? q.v_ld : q.v_n;
            auto iw = wmap.find(kw);
            if (iw == wmap.end()) {
                // W = C op(B)^T   (c_m x v_n), as phase 1 of H*psi
                const bool tb = !(q.conj & 2);
                const uint32_t si = (uint32_t)ch.w.s.size();
                ch.w.s.push_back(make_seg(q.c_off, cld, false, q.b_off,
                                          tb ? q.c_n : q.v_n, tb, q.c_n, 1.0));
                add_tiles(ch.w.t, ch.w_len, q.c_m, q.v_n, q.c_n, si, si + 1, 0);
                iw = wmap.emplace(kw, ch.w_len).first;
                ch.w_len += sz;
            }
            auto iz = zmap.find(kz);
            if (iz == zmap.end()) {
                // Z = op(A)^T V   (c_m x v_n); op(A)^T is A^T for a plain A
                // (stored v_m x c_m) and A for a transposed one
                const bool ta = !(q.conj & 1);
                const uint32_t si = (uint32_t)ch.z.s.size();
                ch.z.s.push_back(make_seg(q.a_off, ta ? q.c_m : q.v_m, ta,
                                          q.v_off, vld, false, q.v_m, 1.0));
                add_tiles(ch.z.t, ch.z_len, q.c_m, q.v_n, q.v_m, si, si + 1, 0);
                iz = zmap.emplace(kz, ch.z_len).first;
                ch.z_len += sz;
            }
            pwz[i1] = {iw->second, iz->second};
        }
        if (i1 == i0) // a pair list of empty pairs only
            i1 = n;
        std::map<uint64_t, std::vector<size_t>> a0, a1, b0, b1;
        for (const auto &kv : pwz) {
            const b2g_pair &q = pairs[kv.first];
            ((q.conj & 1) ? a1 : a0)[q.a_off].push_back(kv.first);
            ((q.conj & 2) ? b1 : b0)[q.b_off].push_back(kv.first);
        }
        // dA = alpha V W^T  (left V from v, right W^T)
        add_pass(ch.l0, a0, pairs,
                 [&](const b2g_pair &q, size_t i, uint32_t &r, uint32_t &c) {
                     r = q.v_m, c = q.c_m;
                     return make_seg(q.v_off, q.v_ld ? q.v_ld : q.v_n, false,
                                     pwz[i].first, q.v_n, true, q.v_n,
                                     q.alpha);
                 });
        // dA = alpha W V^T  (left W from scratch, right V^T)
        add_pass(ch.l1, a1, pairs,
                 [&](const b2g_pair &q, size_t i, uint32_t &r, uint32_t &c) {
                     r = q.c_m, c = q.v_m;
                     return make_seg(pwz[i].first, q.v_n, false, q.v_off,
                                     q.v_ld ? q.v_ld : q.v_n, true, q.v_n,
                                     q.alpha);
                 });
        // dB = alpha Z^T C  (left Z^T from scratch, right C from c)
        add_pass(ch.r0, b0, pairs,
                 [&](const b2g_pair &q, size_t i, uint32_t &r, uint32_t &c) {
                     r = q.v_n, c = q.c_n;
                     return make_seg(pwz[i].second, q.v_n, true, q.c_off,
                                     q.c_ld ? q.c_ld : q.c_n, false, q.c_m,
                                     q.alpha);
                 });
        // dB = alpha C^T Z  (left C^T from c, right Z from scratch)
        add_pass(ch.r1, b1, pairs,
                 [&](const b2g_pair &q, size_t i, uint32_t &r, uint32_t &c) {
                     r = q.c_n, c = q.v_n;
                     return make_seg(q.c_off, q.c_ld ? q.c_ld : q.c_n, true,
                                     pwz[i].second, q.v_n, false, q.c_m,
                                     q.alpha);
                 });
        for (Launch *L : {&ch.w, &ch.z, &ch.l0, &ch.l1, &ch.r0, &ch.r1})
            for (int c = 0; c < kTileConfigs; c++) {
                sort_tasks(L->t[c], L->s);
                g->flops += task_flops(L->t[c], L->s, c);
            }
        const uint64_t llen = p->lbuf->n, rlen = p->rbuf->n;
        const uint64_t xl = p->x_len, yl = p->y_len;
        const bool ok = launch_in_bounds(ch.w, xl, rlen, ch.w_len) &&
                        launch_in_bounds(ch.z, llen, yl, ch.z_len) &&
                        launch_in_bounds(ch.l0, yl, ch.w_len, llen) &&
                        launch_in_bounds(ch.l1, ch.w_len, yl, llen) &&
                        launch_in_bounds(ch.r0, ch.z_len, xl, rlen) &&
                        launch_in_bounds(ch.r1, xl, ch.z_len, rlen);
        if (!ok) {
            delete g;
            set_error("gradient plan: a launch would access outside its "
                      "buffers");
            return B2G_ERR_STATE;
        }
        g->s_max = std::max(g->s_max, ch.w_len + ch.z_len);
        g->chunks.push_back(std::move(ch));
        i0 = i1;
    }
    // upload all task and segment arrays in one allocation
    size_t bytes = 0;
    auto each = [&](auto &&f) {
        for (auto &ch : g->chunks)
            for (Launch *L : {&ch.w, &ch.z, &ch.l0, &ch.l1, &ch.r0, &ch.r1})
                f(*L);
    };
    each([&](Launch &L) {
        for (int c = 0; c < kTileConfigs; c++)
            bytes += (L.t[c].size() * sizeof(Task) + 255) & ~size_t(255);
        bytes += (L.s.size() * sizeof(Seg) + 255) & ~size_t(255);
    });
    std::vector<char> host(std::max<size_t>(bytes, 1));
    size_t off = 0;
    std::vector<size_t> offs;
    each([&](Launch &L) {
        for (int c = 0; c < kTileConfigs; c++) {
            memcpy(host.data() + off, L.t[c].data(), L.t[c].size() * sizeof(Task));
            offs.push_back(off);
            off += (L.t[c].size() * sizeof(Task) + 255) & ~size_t(255);
        }
        memcpy(host.data() + off, L.s.data(), L.s.size() * sizeof(Seg));
        offs.push_back(off);
        off += (L.s.size() * sizeof(Seg) + 255) & ~size_t(255);
    });
    hipError_t e = hipMalloc(&g->dev, std::max<size_t>(bytes, 1));
    if (e == hipSuccess)
        e = hipMemcpy(g->dev, host.data(), std::max<size_t>(bytes, 1),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMalloc(&g->scratch,
                      std::max<uint64_t>(g->s_max, 1) * sizeof(double));
    if (e != hipSuccess) {
        grad_plan_destroy(g);
        set_error(std::string("gradient plan upload: ") + hipGetErrorString(e));
        return B2G_ERR_MEMORY;
    }
    size_t k = 0;
    char *base = (char *)g->dev;
    each([&](Launch &L) {
        for (int c = 0; c < kTileConfigs; c++)
            L.d_t[c] = (Task *)(base + offs[k++]);
        L.d_s = (Seg *)(base + offs[k++]);
    });
    *out = g;
    return B2G_OK;
}

// stage 1 (PH 3) overwrites its scratch; stage 2 (PH 4) adds to the
// gradient buffers
template <int PH>
static hipError_t run_launch(const Launch &L, const double *A, const double *B,
                             double *D, int grid, hipStream_t st) {
    const int beta0 = PH == 4 ? 1 : 0;
    for (int c = 0; c < kTileConfigs; c++) {
        const hipError_t e = launch_seg_gemm<double, double, double, double, PH>(
            c, A, B, D, D, L.d_t[c], (int)L.t[c].size(), L.d_s, 0, 0, 0, 0, 1.0,
            beta0, 1, grid, st);
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

} // namespace b2g

using namespace b2g;

extern "C" int b2g_grad_ops(b2g_ctx ctx, b2g_plan p, const double *v,
                            const double *c, int n_rhs, double *gl,
                            double *gr, int accumulate) {
    B2G_REQUIRE(ctx && p && v && c && gl && gr, B2G_ERR_ARG, "NULL arg");
    B2G_REQUIRE(n_rhs >= 1, B2G_ERR_ARG, "bad n_rhs");
    B2G_HIP(hipSetDevice(ctx->device));
    if (p->grad == nullptr) {
        const int r = grad_plan_build(p, &p->grad);
        if (r != B2G_OK)
            return r;
    }
    const GradPlan *g = p->grad;
    const int grid = ctx->num_cu * 8;
    hipStream_t st = ctx->stream;
    if (!accumulate) {
        B2G_HIP(launch_fill(0.0, gl, p->lbuf->n, st));
        B2G_HIP(launch_fill(0.0, gr, p->rbuf->n, st));
    }
    const double *L = (const double *)p->lbuf->ptr;
    const double *R = (const double *)p->rbuf->ptr;
    for (int r = 0; r < n_rhs; r++) {
        const double *vr = v + (size_t)r * p->y_len;
        const double *cr = c + (size_t)r * p->x_len;
        for (const GradChunk &ch : g->chunks) {
            double *W = g->scratch, *Z = g->scratch + ch.w_len;
            B2G_HIP(run_launch<3>(ch.w, cr, R, W, grid, st));
            B2G_HIP(run_launch<3>(ch.z, L, vr, Z, grid, st));
            B2G_HIP(run_launch<4>(ch.l0, vr, W, gl, grid, st));
            B2G_HIP(run_launch<4>(ch.l1, W, vr, gl, grid, st));
            B2G_HIP(run_launch<4>(ch.r0, Z, cr, gr, grid, st));
            B2G_HIP(run_launch<4>(ch.r1, cr, Z, gr, grid, st));
        }
    }
    return B2G_OK;
}

extern "C" int b2g_grad_flops(b2g_plan p, double *flops) {
    B2G_REQUIRE(p && flops, B2G_ERR_ARG, "NULL arg");
    if (p->grad == nullptr) {
        B2G_HIP(hipSetDevice(p->ctx->device));
        const int r = grad_plan_build(p, &p->grad);
        if (r != B2G_OK)
            return r;
    }
    *flops = p->grad->flops;
    return B2G_OK;
}
