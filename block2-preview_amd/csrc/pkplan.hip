// Split-BF16 replay on packed operand planes: derivation from the replay
// plan, device upload, scratch and operator re-packing.
//
// Every GEMM operand of the split-BF16 kernels is a packed matrix (rows run
// along the contraction index, row length K rounded up to 8, hi and lo BF16
// planes; gemm_bf16x3.hip).  From the plan's chunks:
//   * phase 1 (W = C op(B)): A = the packed wavefunction window C (c_m x c_n,
//     packed per H*psi), B = the packed right-operator block in the
//     orientation op(B)^T (v_n x c_n); the product is written as packed W^T
//     (v_n x round8(c_m));
//   * phase 2 (V += alpha op(A) W): A = the packed left-operator block in the
//     orientation op(A) (v_m x c_m), B = packed W^T; the FP64 output tasks,
//     split-K reductions and partial buffer are the plan's own.
// A block read in both orientations is packed twice; operator planes cost
// 4 bytes per packed element (half of the FP64 buffers they come from).
#include "common.hpp"
#include "launch.hpp"
#include "replay.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <thread>
#include <tuple>
#include <unordered_map>

namespace b2g {

namespace {

uint64_t round8(uint64_t v) { return (v + 7) & ~(uint64_t)7; }

// packed matrices of one source buffer, deduplicated by
// (offset, row stride, orientation, rows, K)
struct Packer {
    typedef std::tuple<uint64_t, uint32_t, uint32_t, uint32_t, uint32_t> Key;
    struct KeyHash {
        size_t operator()(const Key &k) const {
            uint64_t h = std::get<0>(k) * 0x9E3779B97F4A7C15ull;
            h ^= ((uint64_t)std::get<1>(k) << 32 | std::get<2>(k)) +
                 0x632BE59BD9B4E019ull + (h << 6) + (h >> 2);
            h ^= ((uint64_t)std::get<3>(k) << 32 | std::get<4>(k)) + (h << 6) +
                 (h >> 2);
            return (size_t)h;
        }
    };
    FlatMap<Key, uint64_t, KeyHash> where;
    std::vector<PackMat> mats;
    uint64_t len = 0, ngroups = 0; // packed elements, 8-element groups
    uint64_t src_len = 0;

    // packed offset of rows x K matrix (r, k) = src[off + r*ld + k]
    // (trans: src[off + k*ld + r]); -1 on an out-of-range source
    int64_t add(uint64_t off, uint32_t ld, bool trans, uint32_t rows,
                uint32_t K) {
        const Key key(off, ld, trans ? 1u : 0u, rows, K);
        if (const uint64_t *w = where.find(key))
            return (int64_t)*w;
        if (rows == 0 || K == 0)
            return -1;
        const uint64_t last =
            off + (trans ? (uint64_t)(K - 1) * ld + (rows - 1)
                         : (uint64_t)(rows - 1) * ld + (K - 1));
        if (last >= src_len)
            return -1;
        const uint64_t pld = round8(K);
        if ((uint64_t)rows * pld >= ((uint64_t)1 << 28))
            return -1;
        const uint64_t base = len;
        PackMat m;
        m.src = off;
        m.dst = base;
        m.src_ld = ld;
        m.dst_ld = (uint32_t)pld;
        m.rows = rows;
        m.cols = K;
        m.trans = trans ? 1u : 0u;
        m.pad = 0;
        m.g0 = ngroups;
        mats.push_back(m);
        ngroups += (uint64_t)rows * (pld / 8);
        len += (uint64_t)rows * pld;
        where.emplace(key, base);
        return (int64_t)base;
    }
};

// phase-1 tasks grouped by their right-operator block (B2G_PK_GROUP_B=0:
// by W block)
static bool pk_group_b() {
    const char *e = getenv("B2G_PK_GROUP_B");
    return e == nullptr || atoi(e) != 0;
}

struct HostChunk {
    std::vector<Task> t1[kTileConfigs], t2[kTileConfigs];
    std::vector<Seg> s1, s2;
    std::vector<RedTask> red;
    uint64_t p_len = 0;
    XcdSplit x1[kTileConfigs] = {}, x2[kTileConfigs] = {};
};

} // namespace

void pk_destroy(PkPlan *k) {
    if (k == nullptr)
        return;
    if (k->dev)
        hipFree(k->dev);
    if (k->op)
        hipFree(k->op);
    if (k->scratch)
        hipFree(k->scratch);
    delete k;
}

static int pk_pack_ops(b2g_plan p, PkPlan *k) {
    const int grid = p->ctx->num_cu * 8;
    hipStream_t st = p->ctx->stream;
    __bf16 *op = (__bf16 *)k->op;
    B2G_HIP(launch_pack_bf16((const double *)p->lbuf->ptr, op, op + k->l_len,
                             k->d_lmats, k->n_lmats, k->n_lgroups, 0, 0, 1,
                             grid, st));
    B2G_HIP(launch_pack_bf16((const double *)p->rbuf->ptr, op + 2 * k->l_len,
                             op + 2 * k->l_len + k->r_len, k->d_rmats,
                             k->n_rmats, k->n_rgroups, 0, 0, 1, grid, st));
    return B2G_OK;
}

static int pk_build_into(b2g_plan p, PkPlan *k) {
    SetupTimer tm("packed plan");
    tile_env_refresh();
    const bool group_b = pk_group_b();
    Packer L, R, X;
    L.src_len = p->lbuf->n;
    R.src_len = p->rbuf->n;
    X.src_len = p->x_len;
    {
        size_t n1 = 0, n2 = 0;
        for (const Chunk &ch : p->chunks)
            n1 += ch.s1.size(), n2 += ch.s2.size();
        X.where.reserve(n1);
        R.where.reserve(n1);
        L.where.reserve(n2);
        X.mats.reserve(n1);
        R.mats.reserve(n1);
        L.mats.reserve(n2);
    }
    std::vector<HostChunk> hc(p->chunks.size());
    for (size_t ic = 0; ic < p->chunks.size(); ic++) {
        const Chunk &ch = p->chunks[ic];
        HostChunk &h = hc[ic];
        // rows / columns of each segment's output block
        std::vector<uint32_t> m1(ch.s1.size(), 0), n1(ch.s1.size(), 0);
        std::vector<uint64_t> w1(ch.s1.size(), 0);
        for (int c = 0; c < kTileConfigs; c++)
            for (const Task &t : ch.t1[c]) {
                B2G_REQUIRE(t.seg_end == t.seg_begin + 1 &&
                                t.seg_begin < ch.s1.size(),
                            B2G_ERR_STATE, "packed plan: phase-1 task shape");
                m1[t.seg_begin] = t.M, n1[t.seg_begin] = t.N;
                w1[t.seg_begin] = t.d_off;
            }
        std::vector<uint32_t> m2(ch.s2.size(), 0);
        for (const BlockJob &j : ch.jobs2) {
            B2G_REQUIRE(j.sb <= j.se && j.se <= ch.s2.size(), B2G_ERR_STATE,
                        "packed plan: phase-2 block job shape");
            for (uint32_t s = j.sb; s < j.se; s++)
                m2[s] = j.M;
        }
        tm.mark("segment shapes");
        // packed W^T layout of the phase-1 segments, keyed by the plan's W
        // offset (read by the phase-2 pass)
        FlatMap<uint64_t, std::pair<uint64_t, uint32_t>, U64Hash> wmap(
            ch.s1.size());
        std::vector<uint64_t> wpk(ch.s1.size(), 0);
        uint64_t wlen = 0;
        for (size_t si = 0; si < ch.s1.size(); si++) {
            B2G_REQUIRE(m1[si] > 0 && n1[si] > 0, B2G_ERR_STATE,
                        "packed plan: phase-1 segment without a task");
            const uint32_t ldw = (uint32_t)round8(m1[si]);
            B2G_REQUIRE((uint64_t)n1[si] * ldw < ((uint64_t)1 << 28),
                        B2G_ERR_ARG, "packed plan: W block of 2^28 elements");
            wpk[si] = wlen;
            *wmap.emplace(w1[si], std::make_pair(wlen, ldw)).first =
                std::make_pair(wlen, ldw);
            wlen += (uint64_t)n1[si] * ldw;
        }
        // Five independent parts on the host workers: the packed lists of
        // the wavefunction windows, of the right operators (phase 1) and of
        // the left operators (phase 2), and the tiles of either phase.  They
        // share only read-only inputs; each packer is used by one part, so
        // the packed layouts are those of a serial pass.  The tiles depend
        // on the segments' depths only, which packing keeps.
        std::vector<int64_t> xa(ch.s1.size()), rb(ch.s1.size());
        const char *err[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
        std::vector<std::function<void()>> jobs;
        jobs.emplace_back([&]() {
            for (size_t si = 0; si < ch.s1.size(); si++) {
                const Seg &s = ch.s1[si];
                xa[si] = X.add(s.a_off, s.lda, false, m1[si], s.K);
                if (xa[si] < 0) {
                    err[0] = "packed plan: phase-1 operand outside its buffer";
                    return;
                }
            }
        });
        jobs.emplace_back([&]() {
            for (size_t si = 0; si < ch.s1.size(); si++) {
                const Seg &s = ch.s1[si];
                rb[si] = R.add(s.b_off, s.ldb, !(s.flags & 2), n1[si], s.K);
                if (rb[si] < 0) {
                    err[1] = "packed plan: phase-1 operand outside its buffer";
                    return;
                }
            }
        });
        jobs.emplace_back([&]() {
            h.s2.reserve(ch.s2.size());
            for (size_t si = 0; si < ch.s2.size(); si++) {
                const Seg &s = ch.s2[si];
                const std::pair<uint64_t, uint32_t> *wp = wmap.find(s.b_off);
                if (wp == nullptr || wp->second != round8(s.K)) {
                    err[2] = "packed plan: phase-2 W not found";
                    return;
                }
                if (m2[si] == 0) {
                    err[2] = "packed plan: phase-2 segment without a task";
                    return;
                }
                const int64_t a =
                    L.add(s.a_off, s.lda, s.flags & 1, m2[si], s.K);
                if (a < 0) {
                    err[2] = "packed plan: phase-2 operand outside its buffer";
                    return;
                }
                Seg q = s;
                q.a_off = (uint64_t)a, q.lda = (uint32_t)round8(s.K);
                q.b_off = wp->first, q.ldb = wp->second;
                q.flags = 0;
                h.s2.push_back(q);
            }
        });
        // one W^T block per phase-1 segment, tiled with the split-BF16
        // rates (segment K from the plan's own segments)
        jobs.emplace_back([&]() {
            tile_env_refresh();
            for (size_t si = 0; si < ch.s1.size(); si++)
                for (const Tile &t :
                     tile_cover(m1[si], n1[si], ch.s1[si].K, 1))
                    h.t1[cfg_of(t)].push_back(make_task(
                        wpk[si], (uint32_t)round8(m1[si]), m1[si], n1[si], t,
                        (uint32_t)si, (uint32_t)si + 1, 0));
        });
        // output blocks re-tiled likewise, with split-K partial slots and
        // reductions of their own
        jobs.emplace_back([&]() {
            tile_env_refresh();
            for (const BlockJob &j : ch.jobs2)
                add_split_tiles(ch.s2, j, 1, h.t2, h.red, h.p_len);
        });
        run_host_jobs(jobs);
        for (const char *e : err)
            B2G_REQUIRE(e == nullptr, B2G_ERR_ARG, e ? e : "");
        h.s1.reserve(ch.s1.size());
        for (size_t si = 0; si < ch.s1.size(); si++) {
            Seg q = ch.s1[si];
            q.a_off = (uint64_t)xa[si], q.lda = (uint32_t)round8(q.K);
            q.b_off = (uint64_t)rb[si], q.ldb = (uint32_t)round8(q.K);
            q.flags = 0, q.alpha = 1.0;
            h.s1.push_back(q);
        }
        tm.mark("pack lists and tiles (host workers)");
        k->w_max = std::max(k->w_max, wlen);
        k->p_max = std::max(k->p_max, h.p_len);
    }
    k->l_len = L.len, k->r_len = R.len, k->x_len = X.len;
    if (getenv("B2G_TIME_SETUP") != nullptr)
        fprintf(stderr,
                "b2g setup packed plan: %zu/%zu/%zu matrices, %llu/%llu/%llu "
                "packed elements (L/R/psi), sources %zu/%zu\n",
                L.mats.size(), R.mats.size(), X.mats.size(),
                (unsigned long long)L.len, (unsigned long long)R.len,
                (unsigned long long)X.len, (size_t)L.src_len,
                (size_t)R.src_len);
    // one allocation for the pack lists (first: uploaded and used by the
    // operator packing while the host orders the tasks) and every task,
    // segment and reduction array
    const size_t mat_bytes =
        (L.mats.size() + R.mats.size() + X.mats.size()) * sizeof(PackMat);
    size_t bytes = mat_bytes;
    for (auto &h : hc) {
        for (int c = 0; c < kTileConfigs; c++)
            bytes += (h.t1[c].size() + h.t2[c].size()) * sizeof(Task);
        bytes += (h.s1.size() + h.s2.size()) * sizeof(Seg);
        bytes += h.red.size() * sizeof(RedTask);
    }
    B2G_REQUIRE(L.ngroups < (1ull << 40) && R.ngroups < (1ull << 40) &&
                    X.ngroups < (1ull << 40) && L.mats.size() < (1u << 31) &&
                    R.mats.size() < (1u << 31) && X.mats.size() < (1u << 31),
                B2G_ERR_ARG, "packed plan: too many packed matrices");
    char *host = nullptr;
    const int rs = ctx_stage(p->ctx, bytes + 16, &host);
    if (rs != B2G_OK)
        return rs;
    B2G_HIP(hipMalloc(&k->dev, bytes + 16));
    size_t off = 0;
    auto put = [&](const void *src, size_t nb) {
        const size_t o = off;
        if (nb)
            memcpy(host + o, src, nb);
        off += nb;
        return (char *)k->dev + o;
    };
    k->d_lmats = (PackMat *)put(L.mats.data(), L.mats.size() * sizeof(PackMat));
    k->d_rmats = (PackMat *)put(R.mats.data(), R.mats.size() * sizeof(PackMat));
    k->d_xmats = (PackMat *)put(X.mats.data(), X.mats.size() * sizeof(PackMat));
    k->n_lmats = (int)L.mats.size();
    k->n_rmats = (int)R.mats.size();
    k->n_xmats = (int)X.mats.size();
    k->n_lgroups = L.ngroups;
    k->n_rgroups = R.ngroups;
    k->n_xgroups = X.ngroups;
    if (mat_bytes)
        B2G_HIP(hipMemcpyAsync(k->dev, host, mat_bytes, hipMemcpyHostToDevice,
                               p->ctx->stream));
    B2G_HIP(hipMalloc(&k->op, std::max<uint64_t>(2 * (k->l_len + k->r_len), 8) *
                                  sizeof(uint16_t)));
    tm.mark("pack lists uploaded, planes allocated");
    {
        const int r = pk_pack_ops(p, k);
        if (r != B2G_OK) {
            hipStreamSynchronize(p->ctx->stream);
            return r;
        }
    }
    // task order while the device packs the operators: every (chunk,
    // phase, configuration) launch is one item, taken largest first by the
    // host workers
    {
        struct Item {
            size_t n;
            HostChunk *h;
            int phase, c;
        };
        std::vector<Item> items;
        for (auto &h : hc)
            for (int c = 0; c < kTileConfigs; c++) {
                if (!h.t1[c].empty())
                    items.push_back(Item{h.t1[c].size(), &h, 1, c});
                if (!h.t2[c].empty())
                    items.push_back(Item{h.t2[c].size(), &h, 2, c});
            }
        std::stable_sort(items.begin(), items.end(),
                         [](const Item &a, const Item &b) { return a.n > b.n; });
        std::atomic<size_t> next{0};
        // (task order: tiles of one output block (one segment list) share
        // their operand panels and stay on one XCD, back to back, where its
        // L2 serves the re-reads; the reductions address partial slots by
        // task offsets, which reordering keeps)
        auto worker = [&]() {
            for (size_t i = next++; i < items.size(); i = next++) {
                const Item &it = items[i];
                HostChunk &h = *it.h;
                if (it.phase == 1)
                    h.x1[it.c] = xcd_order(h.t1[it.c], h.s1, true, group_b);
                else
                    h.x2[it.c] = xcd_order(h.t2[it.c], h.s2, true);
            }
        };
        const size_t nw = std::min<size_t>(items.size(), 4);
        std::vector<std::function<void()>> jobs(std::max<size_t>(nw, 1),
                                                worker);
        run_host_jobs(jobs);
    }
    tm.mark("task order (host workers)");
    for (auto &h : hc) {
        PkChunk pc;
        for (int c = 0; c < kTileConfigs; c++) {
            pc.d_t1[c] = (Task *)put(h.t1[c].data(),
                                     h.t1[c].size() * sizeof(Task));
            pc.d_t2[c] = (Task *)put(h.t2[c].data(),
                                     h.t2[c].size() * sizeof(Task));
            pc.x1[c] = h.x1[c];
            pc.x2[c] = h.x2[c];
            pc.n1[c] = h.t1[c].size();
            pc.n2[c] = h.t2[c].size();
            pc.f1[c] = task_flops(h.t1[c], h.s1, c);
            pc.f2[c] = task_flops(h.t2[c], h.s2, c);
        }
        pc.d_red = (RedTask *)put(h.red.data(), h.red.size() * sizeof(RedTask));
        pc.n_red = (int)h.red.size();
        pc.d_s1 = (Seg *)put(h.s1.data(), h.s1.size() * sizeof(Seg));
        pc.d_s2 = (Seg *)put(h.s2.data(), h.s2.size() * sizeof(Seg));
        k->chunks.push_back(pc);
    }
    if (bytes > mat_bytes)
        B2G_HIP(hipMemcpyAsync((char *)k->dev + mat_bytes, host + mat_bytes,
                               bytes - mat_bytes, hipMemcpyHostToDevice,
                               p->ctx->stream));
    B2G_HIP(hipStreamSynchronize(p->ctx->stream));
    tm.mark("tasks uploaded, operators packed");
    return B2G_OK;
}

int pk_build(b2g_plan p) {
    if (p->pk != nullptr)
        return B2G_OK;
    PkPlan *k = new PkPlan();
    const int r = pk_build_into(p, k);
    if (r != B2G_OK) {
        hipStreamSynchronize(p->ctx->stream);
        pk_destroy(k);
        return r;
    }
    p->pk = k;
    p->scratch_gen++;
    return B2G_OK;
}

int pk_reserve(b2g_plan p, int n_rhs) {
    PkPlan *k = p->pk;
    if (n_rhs <= k->scratch_rhs)
        return B2G_OK;
    if (k->scratch) {
        B2G_HIP(hipStreamSynchronize(p->ctx->stream));
        B2G_HIP(hipFree(k->scratch));
        k->scratch = nullptr;
        k->scratch_rhs = 0;
    }
    const size_t n = std::max<uint64_t>(2 * (k->x_len + k->w_max), 8) * n_rhs;
    B2G_HIP(hipMalloc(&k->scratch, n * sizeof(uint16_t)));
    k->scratch_rhs = n_rhs;
    p->scratch_gen++;
    return B2G_OK;
}

int pk_refresh(b2g_plan p) {
    if (p->pk == nullptr)
        return B2G_OK;
    return pk_pack_ops(p, p->pk);
}

} // namespace b2g
