// Environment blocking and renormalisation on the device: the window-sum
// kernel (block formation, numerical transforms, fused-form swaps), rotation
// plans on the grouped-GEMM replay and block-product lists (two-site
// wavefunction contraction).
//
// Mirrors:
//   TensorFunctions::left_contract / right_contract   src/core/tensor_functions.hpp:2842-2986
//     -> tensor_product (Prod strings)                 :2184-2288
//     -> OperatorFunctions::tensor_product             src/core/operator_functions.hpp:672-711
//     -> GMatrixFunctions::tensor_product              src/core/matrix_functions.hpp:1269-
//   ConnectionInfo::initialize_tp                      src/core/sparse_matrix.hpp:294-422
//   TensorFunctions::left_rotate / right_rotate        src/core/tensor_functions.hpp:2365-2403
//     -> OperatorFunctions::tensor_rotate              src/core/operator_functions.hpp:175-211
//   TensorFunctions::numerical_transform               src/core/tensor_functions.hpp:2462-2516
//     -> OperatorFunctions::iadd                       src/core/operator_functions.hpp:134-174
//   SparseMatrix::swap_to_fused_left / right           src/core/sparse_matrix.hpp:1823-1968
//   SparseMatrix::contract                             src/core/sparse_matrix.hpp:1777-1801
#include "common.hpp"
#include "envplan.hpp"
#include "launch.hpp"
#include "replay.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>
#include <unordered_map>

using namespace b2g;

namespace b2g {

// ------------------------------------------------------------ window sums
// One workgroup-sized piece of a destination window: elements
// [e0, e0 + kWinChunk) of the window in row-major order.
struct WinChunk {
    uint32_t group, e0;
};
constexpr int kWinPer = 8; // elements per thread and chunk
constexpr uint32_t kWinChunk = 256 * kWinPer;

// Every thread keeps its kWinPer elements of the chunk in registers across
// the window's records (summed in record order) and adds them to C once.
// Windows of different groups are disjoint, so chunks never share an
// element.  A and C may be the same buffer (numerical transforms): the plan
// guarantees that no record reads an element a window writes.
__global__ __launch_bounds__(256) void winsum_kernel(
    const WinGroup *__restrict__ groups, const WinRec *__restrict__ recs,
    const WinChunk *__restrict__ chunks, uint32_t nchunks, const double *A,
    const double *B, double *C) {
    for (uint32_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const WinChunk wc = chunks[ch];
        const WinGroup g = groups[wc.group];
        const uint32_t ne = g.rows * g.cols;
        double acc[kWinPer];
        uint32_t rr[kWinPer], cc[kWinPer];
        bool ok[kWinPer];
#pragma unroll
        for (int k = 0; k < kWinPer; k++) {
            const uint32_t e = wc.e0 + threadIdx.x + 256u * k;
            ok[k] = e < ne && e < wc.e0 + kWinChunk;
            rr[k] = ok[k] ? e / g.cols : 0;
            cc[k] = ok[k] ? e - rr[k] * g.cols : 0;
            acc[k] = 0.0;
        }
        for (uint32_t ir = g.rec_begin; ir < g.rec_end; ir++) {
            const WinRec r = recs[ir];
            const bool ta = r.conj & 1, tb = (r.conj >> 1) & 1;
            // dimensions of op(B)
            const uint32_t bmo = tb ? r.bn : r.bm, bno = tb ? r.bm : r.bn;
            const double *a = A + r.a_off;
            const double *b = B + r.b_off;
            if (bmo == 1 && bno == 1) {
                // window = scale * b * op(A)
                const double s = r.scale * b[0];
#pragma unroll
                for (int k = 0; k < kWinPer; k++)
                    if (ok[k])
                        acc[k] += s * (ta ? a[(uint64_t)cc[k] * r.lda + rr[k]]
                                          : a[(uint64_t)rr[k] * r.lda + cc[k]]);
            } else {
                const uint32_t amo = ta ? r.an : r.am,
                               ano = ta ? r.am : r.an;
                if (amo == 1 && ano == 1) {
                    const double s = r.scale * a[0];
#pragma unroll
                    for (int k = 0; k < kWinPer; k++)
                        if (ok[k])
                            acc[k] +=
                                s * (tb ? b[(uint64_t)cc[k] * r.ldb + rr[k]]
                                        : b[(uint64_t)rr[k] * r.ldb + cc[k]]);
                } else {
#pragma unroll
                    for (int k = 0; k < kWinPer; k++)
                        if (ok[k]) {
                            const uint32_t i = rr[k] / bmo, kb = rr[k] % bmo;
                            const uint32_t j = cc[k] / bno, lb = cc[k] % bno;
                            const double av =
                                ta ? a[(uint64_t)j * r.lda + i]
                                   : a[(uint64_t)i * r.lda + j];
                            const double bv =
                                tb ? b[(uint64_t)lb * r.ldb + kb]
                                   : b[(uint64_t)kb * r.ldb + lb];
                            acc[k] += r.scale * av * bv;
                        }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kWinPer; k++)
            if (ok[k])
                C[g.c_off + (uint64_t)rr[k] * g.ldc + cc[k]] += acc[k];
    }
}

static Symm symm_of(int s) { return s == B2G_SZ ? Symm::SZ : Symm::SU2; }

static StateSet to_states(const b2g_state_info &s) {
    StateSet r;
    if (s.n > 0) {
        r.q.assign(s.quanta, s.quanta + s.n);
        r.n.assign(s.n_states, s.n_states + s.n);
    }
    return r;
}

static int check_states(const b2g_state_info &s, const char *what) {
    B2G_REQUIRE(s.n >= 0 && (s.n == 0 || (s.quanta && s.n_states)),
                B2G_ERR_ARG, std::string(what) + ": bad state info");
    for (int i = 1; i < s.n; i++)
        B2G_REQUIRE(s.quanta[i - 1] < s.quanta[i], B2G_ERR_ARG,
                    std::string(what) + ": quanta not strictly ascending");
    return B2G_OK;
}

static int to_block(const b2g_sparse_info &s, const char *what,
                    BlockInfo &b) {
    B2G_REQUIRE(s.n >= 0 && (s.n == 0 || (s.quanta && s.n_states_bra &&
                                          s.n_states_ket && s.block_shifts)),
                B2G_ERR_ARG, std::string(what) + ": bad sparse info");
    for (int i = 1; i < s.n; i++)
        B2G_REQUIRE(s.quanta[i - 1] < s.quanta[i], B2G_ERR_ARG,
                    std::string(what) + ": quanta not strictly ascending");
    b = BlockInfo();
    b.dq = s.delta_quantum;
    b.is_fermion = s.is_fermion != 0;
    if (s.n > 0) {
        b.quanta.assign(s.quanta, s.quanta + s.n);
        b.nbra.assign(s.n_states_bra, s.n_states_bra + s.n);
        b.nket.assign(s.n_states_ket, s.n_states_ket + s.n);
        b.shift.assign(s.block_shifts, s.block_shifts + s.n);
    }
    return B2G_OK;
}

// an operator set: infos sorted by delta quantum plus the operators
struct OpSide {
    std::vector<BlockInfo> infos;
    std::vector<Operator> ops;
    std::vector<InfoRef> refs;
};

static int to_side(int n_infos, const uint64_t *dq,
                   const b2g_sparse_info *infos, int n_ops,
                   const b2g_operator *ops, const char *what, OpSide &s) {
    B2G_REQUIRE(n_infos >= 0 && n_ops >= 0 &&
                    (n_infos == 0 || (dq && infos)) && (n_ops == 0 || ops),
                B2G_ERR_ARG, std::string(what) + ": bad operator set");
    s.infos.resize(n_infos);
    for (int i = 0; i < n_infos; i++) {
        int r = to_block(infos[i], what, s.infos[i]);
        if (r != B2G_OK)
            return r;
        B2G_REQUIRE(infos[i].delta_quantum == dq[i] &&
                        (i == 0 || dq[i - 1] < dq[i]),
                    B2G_ERR_ARG,
                    std::string(what) + ": infos not keyed by ascending dq");
    }
    for (int i = 0; i < n_infos; i++)
        s.refs.emplace_back(dq[i], &s.infos[i]);
    s.ops.resize(n_ops);
    for (int i = 0; i < n_ops; i++) {
        const b2g_operator &o = ops[i];
        B2G_REQUIRE(o.info >= 0 && o.info < n_infos &&
                        s.infos[o.info].dq == o.delta_quantum,
                    B2G_ERR_ARG, std::string(what) + ": operator info");
        s.ops[i] = Operator{o.delta_quantum, o.factor, o.info, o.data_off};
    }
    return B2G_OK;
}

} // namespace b2g

struct b2g_winplan_s {
    b2g_ctx ctx = nullptr;
    b2g::WinPlan plan;
    std::vector<b2g::WinChunk> chunks;
    std::vector<b2g::Connections> tp; // contraction plans: per C info
    bool host_only = false;
    bool b_is_one = false; // B is the plan's own 1.0 (transforms, swaps)
    void *dev = nullptr;
    size_t dev_len = 0; // bytes of the block (from the context's pool)
    b2g::WinGroup *d_groups = nullptr;
    b2g::WinRec *d_recs = nullptr;
    b2g::WinChunk *d_chunks = nullptr;
    double *d_one = nullptr;
    std::vector<std::pair<uint64_t, uint64_t>> reads, writes; // in-place check
};

namespace b2g {

// A device block of at least `bytes` from the context's pool of released
// window-plan blocks (smallest that fits), else a new allocation.
static int win_block(b2g_ctx ctx, size_t bytes, void **out, size_t *len) {
    auto &pool = ctx->win_pool;
    size_t best = pool.size();
    for (size_t i = 0; i < pool.size(); i++)
        if (pool[i].first >= bytes &&
            (best == pool.size() || pool[i].first < pool[best].first))
            best = i;
    if (best != pool.size()) {
        *len = pool[best].first;
        *out = pool[best].second;
        pool.erase(pool.begin() + best);
        return B2G_OK;
    }
    const size_t mib = (size_t)1 << 20;
    const size_t n = (bytes + mib - 1) / mib * mib;
    hipError_t e = hipMalloc(out, n);
    if (e != hipSuccess && !pool.empty()) {
        // none of the released blocks fits: give them back and retry
        (void)hipGetLastError();
        for (auto &b : pool)
            hipFree(b.second);
        pool.clear();
        e = hipMalloc(out, n);
    }
    B2G_HIP(e);
    *len = n;
    return B2G_OK;
}

// bytes of released window-plan blocks a context keeps for reuse; beyond
// it the smallest are freed (a sweep whose bond dimension grows releases
// plans smaller than every later request)
constexpr size_t kWinPoolCap = (size_t)8 << 30;

static int winplan_finish(b2g_winplan w,
                          WinItems &items) {
    std::string err;
    SetupTimer tm("window plan");
    B2G_REQUIRE(w->plan.finish(items, err), B2G_ERR_ARG, err);
    tm.mark("group");
    if (getenv("B2G_TIME_SETUP") != nullptr)
        fprintf(stderr, "b2g window plan: %zu records, %zu windows\n",
                items.size(), w->plan.groups.size());
    for (uint32_t g = 0; g < w->plan.groups.size(); g++) {
        const uint64_t ne = (uint64_t)w->plan.groups[g].rows *
                            w->plan.groups[g].cols;
        B2G_REQUIRE(ne < (1ull << 31), B2G_ERR_ARG,
                    "window of 2^31 elements or more");
        for (uint64_t e = 0; e < ne; e += kWinChunk)
            w->chunks.push_back(WinChunk{g, (uint32_t)e});
    }
    if (w->host_only)
        return B2G_OK;
    const size_t ng = w->plan.groups.size(), nr = w->plan.recs.size(),
                 nc = w->chunks.size();
    const size_t bytes = ng * sizeof(WinGroup) + nr * sizeof(WinRec) +
                         nc * sizeof(WinChunk) + 64;
    B2G_HIP(hipSetDevice(w->ctx->device));
    {
        const int r = win_block(w->ctx, bytes, &w->dev, &w->dev_len);
        if (r != B2G_OK)
            return r;
    }
    char *p = (char *)w->dev;
    w->d_one = (double *)p;
    w->d_groups = (WinGroup *)(p + 64);
    w->d_recs = (WinRec *)((char *)w->d_groups + ng * sizeof(WinGroup));
    w->d_chunks = (WinChunk *)((char *)w->d_recs + nr * sizeof(WinRec));
    // staged through the context's pinned buffer
    char *host = nullptr;
    {
        const int r = ctx_stage(w->ctx, bytes, &host);
        if (r != B2G_OK)
            return r;
    }
    std::memset(host, 0, 64);
    const double one = 1.0;
    std::memcpy(host, &one, sizeof(double));
    if (ng)
        std::memcpy(host + 64, w->plan.groups.data(), ng * sizeof(WinGroup));
    if (nr)
        parallel_memcpy(host + 64 + ng * sizeof(WinGroup),
                        w->plan.recs.data(), nr * sizeof(WinRec));
    if (nc)
        std::memcpy(host + 64 + ng * sizeof(WinGroup) + nr * sizeof(WinRec),
                    w->chunks.data(), nc * sizeof(WinChunk));
    B2G_HIP(hipMemcpyAsync(w->dev, host, bytes, hipMemcpyHostToDevice,
                           w->ctx->stream));
    B2G_HIP(hipStreamSynchronize(w->ctx->stream));
    tm.mark("upload");
    return B2G_OK;
}

// sub-labels of the contraction's C info `l` (Partition::get_uniq_sub_labels,
// src/dmrg/partition.hpp:368-480, full labels): every string's
// (conj, l.combine(bra, ket)), plus the (vacuum, vacuum) class
static int contract_sub_labels(Symm s, uint64_t l,
                               const std::vector<const b2g_contract_term *> &ts,
                               const OpSide &a, const OpSide &b,
                               std::vector<SubLabel> &out) {
    out.clear();
    for (const b2g_contract_term *t : ts) {
        const uint64_t adq = a.ops[t->a].dq, bdq = b.ops[t->b].dq;
        const uint64_t bra = (t->conj & 1) ? q_neg(s, adq) : adq;
        const uint64_t ket = (t->conj & 2) ? bdq : q_neg(s, bdq);
        const uint64_t p = q_combine(s, l, bra, ket);
        B2G_REQUIRE(p != QINVALID, B2G_ERR_ARG,
                    "contraction string violates the symmetry of its "
                    "operator");
        out.emplace_back((uint8_t)t->conj, p);
    }
    const uint64_t p0 = q_combine(s, l, 0, q_at(s, q_sub(s, l, l), 0));
    if (p0 != QINVALID)
        out.emplace_back((uint8_t)0, p0);
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return B2G_OK;
}

} // namespace b2g

extern "C" int b2g_contract_plan(b2g_ctx ctx, const b2g_contract *c,
                                 unsigned flags, b2g_winplan *out) {
    B2G_REQUIRE(c != nullptr && out != nullptr, B2G_ERR_ARG, "NULL arg");
    const bool host_only = (flags & B2G_EFFHAM_HOST_ONLY) != 0;
    B2G_REQUIRE(host_only || ctx != nullptr, B2G_ERR_ARG, "NULL context");
    B2G_REQUIRE(c->symmetry == B2G_SZ || c->symmetry == B2G_SU2, B2G_ERR_ARG,
                "bad symmetry");
    const Symm s = symm_of(c->symmetry);
    int r;
    for (auto x : {std::make_pair(&c->bra_a, "bra_a"),
                   std::make_pair(&c->ket_a, "ket_a"),
                   std::make_pair(&c->bra_b, "bra_b"),
                   std::make_pair(&c->ket_b, "ket_b"),
                   std::make_pair(&c->bra, "bra"),
                   std::make_pair(&c->ket, "ket")})
        if ((r = check_states(*x.first, x.second)) != B2G_OK)
            return r;
    SetupTimer tm("contract plan");
    OpSide A, B, Cs;
    if ((r = to_side(c->n_a_infos, c->a_info_dq, c->a_infos, c->n_a_ops,
                     c->a_ops, "A operators", A)) != B2G_OK ||
        (r = to_side(c->n_b_infos, c->b_info_dq, c->b_infos, c->n_b_ops,
                     c->b_ops, "B operators", B)) != B2G_OK ||
        (r = to_side(c->n_c_infos, c->c_info_dq, c->c_infos, c->n_c_ops,
                     c->c_ops, "C operators", Cs)) != B2G_OK)
        return r;
    B2G_REQUIRE(c->n_terms >= 0 && (c->n_terms == 0 || c->terms),
                B2G_ERR_ARG, "bad terms");
    std::vector<std::vector<const b2g_contract_term *>> by_info(
        Cs.infos.size());
    for (int i = 0; i < c->n_terms; i++) {
        const b2g_contract_term &t = c->terms[i];
        B2G_REQUIRE(t.c >= 0 && t.c < c->n_c_ops && t.a >= 0 &&
                        t.a < c->n_a_ops && t.b >= 0 && t.b < c->n_b_ops &&
                        t.conj < 4,
                    B2G_ERR_ARG, "contraction term out of range");
        by_info[Cs.ops[t.c].info].push_back(&t);
    }
    const StateSet bra_a = to_states(c->bra_a), ket_a = to_states(c->ket_a),
                   bra_b = to_states(c->bra_b), ket_b = to_states(c->ket_b),
                   bra = to_states(c->bra), ket = to_states(c->ket);
    StateConn bc, kc;
    std::string err;
    B2G_REQUIRE(state_connections(s, bra_a, bra_b, bra, bc, err) &&
                    state_connections(s, ket_a, ket_b, ket, kc, err),
                B2G_ERR_ARG, err);
    std::unique_ptr<b2g_winplan_s> w(new b2g_winplan_s());
    w->ctx = ctx;
    w->host_only = host_only;
    SU2Coupling cg;
    w->tp.resize(Cs.infos.size());
    for (size_t i = 0; i < Cs.infos.size(); i++) {
        std::vector<SubLabel> subl;
        if ((r = contract_sub_labels(s, Cs.infos[i].dq, by_info[i], A, B,
                                     subl)) != B2G_OK)
            return r;
        build_tp_connections(w->tp[i], s, cg, Cs.infos[i].dq, subl, bra, ket,
                             bra_a, bra_b, ket_a, ket_b, bc, kc, A.refs,
                             B.refs, Cs.infos[i]);
    }
    tm.mark("connections");
    // records: every string, every connection of its class
    // (OperatorFunctions::tensor_product).  Each term's class range is
    // found here; the records are then written on the host workers into
    // their slots of one array, in term order.
    const int nt = c->n_terms;
    std::vector<int> cls_b(nt, 0), cls_e(nt, 0);
    std::vector<uint64_t> at(nt + 1, 0);
    for (int it = 0; it < nt; it++) {
        const b2g_contract_term &t = c->terms[it];
        const Operator &co = Cs.ops[t.c], &ao = A.ops[t.a], &bo = B.ops[t.b];
        const double scale = t.factor * ao.factor * bo.factor;
        at[it + 1] = at[it];
        if (std::abs(scale) < 1E-20)
            continue;
        const BlockInfo &ci = Cs.infos[co.info];
        const uint64_t abdq = q_combine(
            s, ci.dq, (t.conj & 1) ? q_neg(s, ao.dq) : ao.dq,
            (t.conj & 2) ? bo.dq : q_neg(s, bo.dq));
        int ib = 0, ie = 0;
        B2G_REQUIRE(abdq != QINVALID &&
                        w->tp[co.info].find_class((uint8_t)t.conj, abdq, ib,
                                                  ie),
                    B2G_ERR_ARG, "contraction string without its class");
        cls_b[it] = ib, cls_e[it] = ie;
        at[it + 1] += (uint64_t)(ie - ib);
    }
    WinItems items(at[nt]);
    std::atomic<int> outside{0};
    auto fill = [&](int t0, int t1) {
        for (int it = t0; it < t1; it++) {
            const b2g_contract_term &t = c->terms[it];
            const Operator &co = Cs.ops[t.c], &ao = A.ops[t.a],
                           &bo = B.ops[t.b];
            const BlockInfo &ci = Cs.infos[co.info], &ai = A.infos[ao.info],
                            &bi = B.infos[bo.info];
            const double scale = t.factor * ao.factor * bo.factor;
            const Connections &cn = w->tp[co.info];
            WinItem *out = items.data() + at[it];
            for (int il = cls_b[it]; il < cls_e[it]; il++) {
                const uint32_t xa = cn.ia[il], xb = cn.ib[il],
                               xc = cn.ic[il];
                WinRec rec{};
                rec.a_off = ao.data_off + ai.shift[xa];
                rec.am = ai.nbra[xa], rec.an = ai.nket[xa], rec.lda = rec.an;
                rec.b_off = bo.data_off + bi.shift[xb];
                rec.bm = bi.nbra[xb], rec.bn = bi.nket[xb], rec.ldb = rec.bn;
                rec.conj = t.conj;
                rec.scale = scale * cn.factor[il];
                const uint32_t amo = (t.conj & 1) ? rec.an : rec.am,
                               ano = (t.conj & 1) ? rec.am : rec.an;
                const uint32_t bmo = (t.conj & 2) ? rec.bn : rec.bm,
                               bno = (t.conj & 2) ? rec.bm : rec.bn;
                WinGroup g{};
                g.c_off = co.data_off + ci.shift[xc] + cn.stride[il];
                g.ldc = ci.nket[xc];
                g.rows = amo * bmo;
                g.cols = ano * bno;
                const uint64_t sr = cn.stride[il] / ci.nket[xc],
                               sc = cn.stride[il] % ci.nket[xc];
                if (sr + g.rows > ci.nbra[xc] || sc + g.cols > ci.nket[xc])
                    outside = 1;
                *out++ = WinItem(g, rec);
            }
        }
    };
    {
        // terms split into jobs of about equal record counts
        const size_t nj = plan_jobs(at[nt], 1u << 20);
        std::vector<std::function<void()>> jobs;
        int t0 = 0;
        for (size_t j = 0; j < nj; j++) {
            const uint64_t target = at[nt] * (j + 1) / nj;
            int t1 = t0;
            while (t1 < nt && (j + 1 == nj || at[t1] < target))
                t1++;
            jobs.emplace_back([&fill, t0, t1]() { fill(t0, t1); });
            t0 = t1;
        }
        run_host_jobs(jobs);
    }
    B2G_REQUIRE(!outside, B2G_ERR_ARG, "tensor-product window outside its block");
    tm.mark("records");
    if ((r = winplan_finish(w.get(), items)) != B2G_OK)
        return r;
    *out = w.release();
    return B2G_OK;
}

extern "C" int b2g_winplan_connections(b2g_winplan w, int info, int32_t *n,
                                       uint64_t *n_classes, uint64_t *quanta,
                                       uint32_t *idx, uint64_t *n_conn,
                                       uint32_t *ia, uint32_t *ib,
                                       uint32_t *ic, uint64_t *stride,
                                       double *factor) {
    B2G_REQUIRE(w && info >= 0 && info < (int)w->tp.size(), B2G_ERR_ARG,
                "bad plan or info index");
    const Connections &c = w->tp[info];
    if (n)
        for (int i = 0; i < 5; i++)
            n[i] = c.n[i];
    if (n_classes)
        *n_classes = c.quanta.size();
    if (n_conn)
        *n_conn = c.ia.size();
    if (quanta)
        std::copy(c.quanta.begin(), c.quanta.end(), quanta);
    if (idx)
        std::copy(c.idx.begin(), c.idx.end(), idx);
    if (ia)
        std::copy(c.ia.begin(), c.ia.end(), ia);
    if (ib)
        std::copy(c.ib.begin(), c.ib.end(), ib);
    if (ic)
        std::copy(c.ic.begin(), c.ic.end(), ic);
    if (stride)
        std::copy(c.stride.begin(), c.stride.end(), stride);
    if (factor)
        std::copy(c.factor.begin(), c.factor.end(), factor);
    return B2G_OK;
}

extern "C" int b2g_transform_plan(b2g_ctx ctx, int symmetry, int n_infos,
                                  const uint64_t *info_dq,
                                  const b2g_sparse_info *infos, int n_ops,
                                  const b2g_operator *ops, int n_terms,
                                  const b2g_transform_term *terms,
                                  unsigned flags, b2g_winplan *out) {
    B2G_REQUIRE(out != nullptr, B2G_ERR_ARG, "NULL arg");
    const bool host_only = (flags & B2G_EFFHAM_HOST_ONLY) != 0;
    B2G_REQUIRE(host_only || ctx != nullptr, B2G_ERR_ARG, "NULL context");
    B2G_REQUIRE(symmetry == B2G_SZ || symmetry == B2G_SU2, B2G_ERR_ARG,
                "bad symmetry");
    B2G_REQUIRE(n_terms >= 0 && (n_terms == 0 || terms), B2G_ERR_ARG,
                "bad terms");
    const Symm s = symm_of(symmetry);
    SetupTimer tm("transform plan");
    OpSide O;
    int r;
    if ((r = to_side(n_infos, info_dq, infos, n_ops, ops, "operators", O)) !=
        B2G_OK)
        return r;
    tm.mark("operators");
    SU2Coupling cg;
    std::unique_ptr<b2g_winplan_s> w(new b2g_winplan_s());
    w->ctx = ctx;
    w->host_only = host_only;
    w->b_is_one = true;
    WinItems items;
    std::vector<uint8_t> written(n_ops, 0), read(n_ops, 0);
    // per (c info, a info, conj): the a block of each c block (-1: none)
    // and its transpose factor, shared by every term with those infos
    struct BlockMap {
        std::vector<int> xa;
        std::vector<long double> cgf;
    };
    std::map<std::tuple<int, int, int>, BlockMap> maps;
    auto block_map = [&](const Operator &co, const Operator &ao,
                         int conj) -> const BlockMap & {
        auto key = std::make_tuple(co.info, ao.info, conj);
        auto it = maps.find(key);
        if (it != maps.end())
            return it->second;
        const BlockInfo &ci = O.infos[co.info], &ai = O.infos[ao.info];
        BlockMap m;
        m.xa.assign(ci.n(), -1);
        m.cgf.assign(ci.n(), 1.0L);
        for (int xc = 0; xc < ci.n(); xc++) {
            const uint64_t bra = q_get_bra(s, ci.quanta[xc], ci.dq);
            const uint64_t ket = q_get_ket(s, ci.quanta[xc]);
            const uint64_t aq = conj ? q_combine(s, ao.dq, ket, bra)
                                     : q_combine(s, ao.dq, bra, ket);
            if (aq == QINVALID)
                continue;
            m.xa[xc] = ai.find_state(aq);
            if (m.xa[xc] >= 0 && conj && s == Symm::SU2)
                m.cgf[xc] = cg.transpose(q_twos(s, ao.dq), q_twos(s, bra),
                                         q_twos(s, ket));
        }
        return maps.emplace(key, std::move(m)).first->second;
    };
    // terms by destination operator (destinations in first-appearance
    // order, each one's terms in term order), so that the records of a
    // window come out consecutive: block by block of the destination, the
    // window's terms in term order (OperatorFunctions::iadd(c, a, factor,
    // conj) per term)
    std::vector<int> dest_order, dest_of(n_ops, -1);
    std::vector<std::vector<int>> by_dest;
    for (int it = 0; it < n_terms; it++) {
        const b2g_transform_term &t = terms[it];
        B2G_REQUIRE(t.c >= 0 && t.c < n_ops && t.a >= 0 && t.a < n_ops &&
                        t.conj < 2,
                    B2G_ERR_ARG, "transform term out of range");
        written[t.c] = 1;
        read[t.a] = 1;
        if (std::abs(O.ops[t.a].factor * t.factor) < 1E-20)
            continue;
        if (dest_of[t.c] < 0) {
            dest_of[t.c] = (int)by_dest.size();
            dest_order.push_back(t.c);
            by_dest.emplace_back();
        }
        by_dest[dest_of[t.c]].push_back(it);
    }
    tm.mark("destinations");
    // Per destination: whether its terms sum as whole operators, its block
    // maps (built here, on this thread) and its record count; then the
    // records of the destinations are written on the host workers into
    // their slots of one array, in the serial order.
    const size_t nd = dest_order.size();
    std::vector<uint8_t> whole(nd, 0);
    std::vector<std::vector<const BlockMap *>> bms(nd);
    std::vector<uint64_t> at(nd + 1, 0);
    for (size_t d = 0; d < nd; d++) {
        const Operator &co = O.ops[dest_order[d]];
        const BlockInfo &ci = O.infos[co.info];
        // every term of this destination untransposed and of its layout:
        // the sum is one axpy over the whole operator (the reference's
        // iadd does the same when the infos agree), one record per term
        bool wh = ci.total() < (1ull << 31);
        for (int it : by_dest[d])
            wh = wh && terms[it].conj == 0 && O.ops[terms[it].a].info == co.info;
        whole[d] = wh;
        uint64_t n = 0;
        if (wh) {
            n = ci.total() != 0 ? by_dest[d].size() : 0;
        } else {
            bms[d].resize(by_dest[d].size());
            for (size_t q = 0; q < by_dest[d].size(); q++) {
                const b2g_transform_term &t = terms[by_dest[d][q]];
                bms[d][q] = &block_map(co, O.ops[t.a], t.conj);
            }
            for (int xc = 0; xc < ci.n(); xc++)
                for (size_t q = 0; q < by_dest[d].size(); q++)
                    n += bms[d][q]->xa[xc] >= 0;
        }
        at[d + 1] = at[d] + n;
    }
    tm.mark("block maps and counts");
    items.resize(at[nd]);
    std::atomic<const char *> shape_err{nullptr};
    auto fill = [&](size_t d0, size_t d1) {
        for (size_t d = d0; d < d1; d++) {
            const Operator &co = O.ops[dest_order[d]];
            const BlockInfo &ci = O.infos[co.info];
            WinItem *out = items.data() + at[d];
            if (whole[d]) {
                const uint32_t len = (uint32_t)ci.total();
                if (len == 0)
                    continue;
                for (int it : by_dest[d]) {
                    const b2g_transform_term &t = terms[it];
                    const Operator &ao = O.ops[t.a];
                    WinRec rec{};
                    rec.a_off = ao.data_off;
                    rec.am = 1, rec.an = len, rec.lda = len;
                    rec.b_off = 0, rec.bm = rec.bn = rec.ldb = 1;
                    rec.conj = 0;
                    rec.scale = t.factor * ao.factor;
                    WinGroup g{};
                    g.c_off = co.data_off;
                    g.ldc = len;
                    g.rows = rec.am, g.cols = len;
                    *out++ = WinItem(g, rec);
                }
                continue;
            }
            const std::vector<const BlockMap *> &bm = bms[d];
            for (int xc = 0; xc < ci.n(); xc++)
                for (size_t q = 0; q < by_dest[d].size(); q++) {
                    const int xa = bm[q]->xa[xc];
                    if (xa < 0)
                        continue;
                    const b2g_transform_term &t = terms[by_dest[d][q]];
                    const Operator &ao = O.ops[t.a];
                    const BlockInfo &ai = O.infos[ao.info];
                    double f = t.factor * ao.factor;
                    if (t.conj && s == Symm::SU2)
                        f = (double)((long double)f * bm[q]->cgf[xc]);
                    WinRec rec{};
                    rec.a_off = ao.data_off + ai.shift[xa];
                    rec.am = ai.nbra[xa], rec.an = ai.nket[xa];
                    rec.lda = rec.an;
                    rec.b_off = 0, rec.bm = rec.bn = rec.ldb = 1;
                    rec.conj = t.conj;
                    rec.scale = f;
                    WinGroup g{};
                    g.c_off = co.data_off + ci.shift[xc];
                    g.ldc = ci.nket[xc];
                    g.rows = t.conj ? rec.an : rec.am;
                    g.cols = t.conj ? rec.am : rec.an;
                    if (g.rows != ci.nbra[xc] || g.cols != ci.nket[xc])
                        shape_err = "transform block shapes differ";
                    *out++ = WinItem(g, rec);
                }
        }
    };
    {
        // destinations split into jobs of about equal record counts
        const size_t nj = plan_jobs(at[nd], 1u << 20);
        std::vector<std::function<void()>> jobs;
        size_t d0 = 0;
        for (size_t j = 0; j < nj; j++) {
            const uint64_t target = at[nd] * (j + 1) / nj;
            size_t d1 = d0;
            while (d1 < nd && (j + 1 == nj || at[d1] < target))
                d1++;
            jobs.emplace_back([&fill, d0, d1]() { fill(d0, d1); });
            d0 = d1;
        }
        run_host_jobs(jobs);
    }
    const char *serr = shape_err.load();
    B2G_REQUIRE(serr == nullptr, B2G_ERR_ARG, serr ? serr : "");
    for (int i = 0; i < n_ops; i++)
        B2G_REQUIRE(!(written[i] && read[i]), B2G_ERR_ARG,
                    "a transform reads an operator it writes");
    tm.mark("records");
    if ((r = winplan_finish(w.get(), items)) != B2G_OK)
        return r;
    *out = w.release();
    return B2G_OK;
}

extern "C" int b2g_swap_plan(b2g_ctx ctx, int symmetry, int to_left,
                             const b2g_state_info *l, const b2g_state_info *m,
                             const b2g_state_info *r,
                             const b2g_state_info *old_fused,
                             const b2g_state_info *new_fused,
                             const b2g_sparse_info *old_wfn,
                             const b2g_sparse_info *new_wfn, unsigned flags,
                             b2g_winplan *out) {
    B2G_REQUIRE(l && m && r && old_fused && new_fused && old_wfn && new_wfn &&
                    out,
                B2G_ERR_ARG, "NULL arg");
    const bool host_only = (flags & B2G_EFFHAM_HOST_ONLY) != 0;
    B2G_REQUIRE(host_only || ctx != nullptr, B2G_ERR_ARG, "NULL context");
    B2G_REQUIRE(symmetry == B2G_SZ || symmetry == B2G_SU2, B2G_ERR_ARG,
                "bad symmetry");
    const Symm s = symm_of(symmetry);
    int rc;
    for (auto x : {std::make_pair(l, "l"), std::make_pair(m, "m"),
                   std::make_pair(r, "r"), std::make_pair(old_fused, "old"),
                   std::make_pair(new_fused, "new")})
        if ((rc = check_states(*x.first, x.second)) != B2G_OK)
            return rc;
    BlockInfo ow, nw;
    if ((rc = to_block(*old_wfn, "old wavefunction", ow)) != B2G_OK ||
        (rc = to_block(*new_wfn, "new wavefunction", nw)) != B2G_OK)
        return rc;
    const StateSet L = to_states(*l), M = to_states(*m), R = to_states(*r),
                   OF = to_states(*old_fused), NF = to_states(*new_fused);
    StateConn oc, nc;
    std::string err;
    if (to_left)
        B2G_REQUIRE(state_connections(s, M, R, OF, oc, err) &&
                        state_connections(s, L, M, NF, nc, err),
                    B2G_ERR_ARG, err);
    else
        B2G_REQUIRE(state_connections(s, L, M, OF, oc, err) &&
                        state_connections(s, M, R, NF, nc, err),
                    B2G_ERR_ARG, err);
    SU2Coupling cg;
    auto racah = [&](uint64_t a, uint64_t b, uint64_t c, uint64_t d,
                     uint64_t e, uint64_t f) -> long double {
        if (s == Symm::SZ)
            return 1.0L;
        const int ta = q_twos(s, a), tb = q_twos(s, b), tc = q_twos(s, c),
                  td = q_twos(s, d), te = q_twos(s, e), tf = q_twos(s, f);
        return (long double)(1 - ((ta + tb + tc + td) & 2)) *
               cg.sixj(ta, tb, te, td, tc, tf);
    };
    std::unique_ptr<b2g_winplan_s> w(new b2g_winplan_s());
    w->ctx = ctx;
    w->host_only = host_only;
    w->b_is_one = true;
    WinItems items;
    if (to_left) {
        // [l x (m r)] -> [(l m) x r]
        std::map<std::pair<uint32_t, uint32_t>,
                 std::map<uint32_t, std::vector<std::pair<uint64_t, int>>>>
            mp;
        for (int i = 0; i < ow.n(); i++) {
            const uint64_t bra = q_get_bra(s, ow.quanta[i], ow.dq);
            const uint64_t ket = q_neg(s, q_get_ket(s, ow.quanta[i]));
            const int ib = L.find(bra), ik = OF.find(ket);
            B2G_REQUIRE(ib >= 0 && ik >= 0, B2G_ERR_ARG,
                        "old wavefunction block outside l x (m r)");
            uint64_t p = ow.shift[i];
            for (uint32_t kk = oc.acc[ik]; kk < oc.acc[ik + 1]; kk++) {
                const uint32_t ka = oc.ij[kk].first, kb = oc.ij[kk].second;
                mp[std::make_pair((uint32_t)ib, ka)][kb].emplace_back(p, ik);
                p += (uint64_t)M.n[ka] * R.n[kb];
            }
        }
        for (int i = 0; i < nw.n(); i++) {
            const uint64_t bra = q_get_bra(s, nw.quanta[i], nw.dq);
            const uint64_t ket = q_neg(s, q_get_ket(s, nw.quanta[i]));
            const int ib = NF.find(bra), ik = R.find(ket);
            B2G_REQUIRE(ib >= 0 && ik >= 0, B2G_ERR_ARG,
                        "new wavefunction block outside (l m) x r");
            uint64_t ptr = nw.shift[i];
            for (uint32_t bb = nc.acc[ib]; bb < nc.acc[ib + 1]; bb++) {
                const uint32_t ba = nc.ij[bb].first, bm = nc.ij[bb].second;
                const uint64_t lp = (uint64_t)M.n[bm] * R.n[ik];
                auto it = mp.find(nc.ij[bb]);
                if (it != mp.end()) {
                    auto jt = it->second.find((uint32_t)ik);
                    if (jt != it->second.end())
                        for (auto &t : jt->second) {
                            const uint64_t ket_mr = OF.q[t.second];
                            const double f = (double)(
                                racah(L.q[ba], M.q[bm], nw.dq, ket, bra,
                                      ket_mr) *
                                (long double)std::sqrt(
                                    1.0 * q_multiplicity(s, bra) *
                                    q_multiplicity(s, ket_mr)));
                            WinRec rec{};
                            rec.a_off = t.first;
                            rec.am = L.n[ba], rec.an = (uint32_t)lp;
                            rec.lda = OF.n[t.second];
                            rec.b_off = 0, rec.bm = rec.bn = rec.ldb = 1;
                            rec.scale = f;
                            WinGroup g{};
                            g.c_off = ptr;
                            g.ldc = (uint32_t)lp;
                            g.rows = L.n[ba], g.cols = (uint32_t)lp;
                            items.emplace_back(g, rec);
                        }
                }
                ptr += (uint64_t)L.n[ba] * lp;
            }
        }
    } else {
        // [(l m) x r] -> [l x (m r)]
        std::map<std::pair<uint32_t, uint32_t>,
                 std::map<uint32_t,
                          std::vector<std::tuple<uint64_t, uint64_t, int>>>>
            mp;
        for (int i = 0; i < ow.n(); i++) {
            const uint64_t bra = q_get_bra(s, ow.quanta[i], ow.dq);
            const uint64_t ket = q_neg(s, q_get_ket(s, ow.quanta[i]));
            const int ib = OF.find(bra), ik = R.find(ket);
            B2G_REQUIRE(ib >= 0 && ik >= 0, B2G_ERR_ARG,
                        "old wavefunction block outside (l m) x r");
            uint64_t p = ow.shift[i];
            for (uint32_t bb = oc.acc[ib]; bb < oc.acc[ib + 1]; bb++) {
                const uint32_t ba = oc.ij[bb].first, bm = oc.ij[bb].second;
                const uint64_t lp = (uint64_t)M.n[bm] * R.n[ik];
                mp[std::make_pair(bm, (uint32_t)ik)][ba].emplace_back(p, lp,
                                                                      ib);
                p += (uint64_t)L.n[ba] * lp;
            }
        }
        for (int i = 0; i < nw.n(); i++) {
            const uint64_t bra = q_get_bra(s, nw.quanta[i], nw.dq);
            const uint64_t ket = q_neg(s, q_get_ket(s, nw.quanta[i]));
            const int ib = L.find(bra), ik = NF.find(ket);
            B2G_REQUIRE(ib >= 0 && ik >= 0, B2G_ERR_ARG,
                        "new wavefunction block outside l x (m r)");
            uint64_t ptr = nw.shift[i];
            const uint64_t lp = NF.n[ik];
            for (uint32_t kk = nc.acc[ik]; kk < nc.acc[ik + 1]; kk++) {
                const uint32_t ka = nc.ij[kk].first, kb = nc.ij[kk].second;
                auto it = mp.find(nc.ij[kk]);
                if (it != mp.end()) {
                    auto jt = it->second.find((uint32_t)ib);
                    if (jt != it->second.end())
                        for (auto &t : jt->second) {
                            const uint64_t bra_lm = OF.q[std::get<2>(t)];
                            const double f = (double)(
                                racah(R.q[kb], M.q[ka], nw.dq, bra, ket,
                                      bra_lm) *
                                (long double)std::sqrt(
                                    1.0 * q_multiplicity(s, ket) *
                                    q_multiplicity(s, bra_lm)));
                            WinRec rec{};
                            rec.a_off = std::get<0>(t);
                            rec.am = L.n[ib];
                            rec.an = (uint32_t)std::get<1>(t);
                            rec.lda = rec.an;
                            rec.b_off = 0, rec.bm = rec.bn = rec.ldb = 1;
                            rec.scale = f;
                            WinGroup g{};
                            g.c_off = ptr;
                            g.ldc = (uint32_t)lp;
                            g.rows = L.n[ib], g.cols = rec.an;
                            items.emplace_back(g, rec);
                        }
                }
                ptr += (uint64_t)M.n[ka] * R.n[kb];
            }
        }
    }
    if ((rc = winplan_finish(w.get(), items)) != B2G_OK)
        return rc;
    *out = w.release();
    return B2G_OK;
}

extern "C" int b2g_winplan_run(b2g_winplan w, const double *a,
                               const double *b, double *c, int accumulate) {
    B2G_REQUIRE(w != nullptr, B2G_ERR_ARG, "NULL plan");
    B2G_REQUIRE(!w->host_only, B2G_ERR_STATE, "host-only plan");
    B2G_REQUIRE(c != nullptr && (a != nullptr || w->plan.recs.empty()) &&
                    (b != nullptr || w->b_is_one || w->plan.recs.empty()),
                B2G_ERR_ARG, "NULL data");
    B2G_HIP(hipSetDevice(w->ctx->device));
    hipStream_t st = w->ctx->stream;
    if (!accumulate && w->plan.c_len)
        B2G_HIP(launch_fill(0.0, c, w->plan.c_len, st));
    if (w->chunks.empty())
        return B2G_OK;
    const uint32_t nch = (uint32_t)w->chunks.size();
    const unsigned grid =
        std::min<unsigned>(nch, (unsigned)std::max(1, w->ctx->num_cu * 8));
    hipLaunchKernelGGL(winsum_kernel, dim3(grid), dim3(256), 0, st,
                       w->d_groups, w->d_recs, w->d_chunks, nch, a,
                       w->b_is_one ? w->d_one : b, c);
    B2G_HIP(hipGetLastError());
    return B2G_OK;
}

extern "C" int b2g_winplan_stats(b2g_winplan w, uint64_t *n_groups,
                                 uint64_t *n_recs, uint64_t *src_elems,
                                 uint64_t *dst_elems, uint64_t *c_len) {
    B2G_REQUIRE(w != nullptr, B2G_ERR_ARG, "NULL plan");
    if (n_groups)
        *n_groups = w->plan.groups.size();
    if (n_recs)
        *n_recs = w->plan.recs.size();
    if (src_elems)
        *src_elems = w->plan.src_elems;
    if (dst_elems)
        *dst_elems = w->plan.dst_elems;
    if (c_len)
        *c_len = w->plan.c_len;
    return B2G_OK;
}

extern "C" int b2g_winplan_destroy(b2g_winplan w) {
    if (w == nullptr)
        return B2G_OK;
    if (w->dev) {
        // the plan's launches are done before its block serves another
        hipSetDevice(w->ctx->device);
        hipStreamSynchronize(w->ctx->stream);
        auto &pool = w->ctx->win_pool;
        pool.emplace_back(w->dev_len, w->dev);
        size_t total = 0;
        for (const auto &b : pool)
            total += b.first;
        while (total > kWinPoolCap) {
            auto it = std::min_element(pool.begin(), pool.end());
            total -= it->first;
            hipFree(it->second);
            pool.erase(it);
        }
    }
    delete w;
    return B2G_OK;
}

// ----------------------------------------------------------------- rotation
extern "C" int b2g_rotate_plan(b2g_ctx ctx, int symmetry, int trans,
                               const b2g_sparse_info *rot, b2g_buf rot_buf,
                               int n_a_infos, const uint64_t *a_info_dq,
                               const b2g_sparse_info *a_infos, int n_c_infos,
                               const uint64_t *c_info_dq,
                               const b2g_sparse_info *c_infos, int n_ops,
                               const b2g_operator *a_ops,
                               const b2g_operator *c_ops, size_t a_len,
                               size_t c_len, unsigned flags, b2g_plan *out) {
    B2G_REQUIRE(ctx && rot && rot_buf && out, B2G_ERR_ARG, "NULL arg");
    B2G_REQUIRE(symmetry == B2G_SZ || symmetry == B2G_SU2, B2G_ERR_ARG,
                "bad symmetry");
    const Symm s = symm_of(symmetry);
    BlockInfo R;
    OpSide A, Cs;
    int r;
    if ((r = to_block(*rot, "rotation", R)) != B2G_OK ||
        (r = to_side(n_a_infos, a_info_dq, a_infos, n_ops, a_ops,
                     "A operators", A)) != B2G_OK ||
        (r = to_side(n_c_infos, c_info_dq, c_infos, n_ops, c_ops,
                     "C operators", Cs)) != B2G_OK)
        return r;
    B2G_REQUIRE(R.total() <= rot_buf->n, B2G_ERR_ARG,
                "rotation data shorter than its info");
    std::vector<b2g_pair> pairs;
    std::vector<b2g_block> yb;
    for (int k = 0; k < n_ops; k++) {
        const Operator &ao = A.ops[k], &co = Cs.ops[k];
        const BlockInfo &ai = A.infos[ao.info], &ci = Cs.infos[co.info];
        B2G_REQUIRE(ao.dq == co.dq, B2G_ERR_ARG,
                    "rotated operator changes its delta quantum");
        B2G_REQUIRE(co.data_off + ci.total() <= c_len &&
                        ao.data_off + ai.total() <= a_len,
                    B2G_ERR_ARG, "operator data outside its buffer");
        for (int xc = 0; xc < ci.n(); xc++) {
            b2g_block b;
            b.off = co.data_off + ci.shift[xc];
            b.m = ci.nbra[xc], b.n = ci.nket[xc];
            yb.push_back(b);
        }
        // OperatorFunctions::tensor_rotate
        const double scale = ao.factor;
        if (std::abs(scale) < 1E-20)
            continue;
        for (int xc = 0; xc < ci.n(); xc++) {
            const int xa = ai.find_state(ci.quanta[xc]);
            if (xa < 0)
                continue;
            const uint64_t cq = q_get_bra(s, ci.quanta[xc], ci.dq);
            const uint64_t cqp = q_get_ket(s, ci.quanta[xc]);
            const int ib = R.find_state(cq), ik = R.find_state(cqp);
            if (ib < 0 || ik < 0)
                continue;
            b2g_pair p{};
            p.a_off = R.shift[ib];
            p.b_off = R.shift[ik];
            p.c_off = ao.data_off + ai.shift[xa];
            p.v_off = co.data_off + ci.shift[xc];
            p.c_m = ai.nbra[xa], p.c_n = ai.nket[xa];
            p.v_m = ci.nbra[xc], p.v_n = ci.nket[xc];
            p.conj = trans ? 0 : 3;
            p.kind = 0;
            p.alpha = scale;
            // the rotation blocks must match: left [fused x bond], right
            // [bond x fused]
            const uint32_t bm = trans ? p.v_m : p.c_m, bn = trans ? p.c_m : p.v_m;
            const uint32_t km = trans ? p.v_n : p.c_n, kn = trans ? p.c_n : p.v_n;
            B2G_REQUIRE(R.nbra[ib] == bm && R.nket[ib] == bn &&
                            R.nbra[ik] == km && R.nket[ik] == kn,
                        B2G_ERR_ARG, "rotation block shapes differ");
            pairs.push_back(p);
        }
    }
    return b2g_upload_plan(ctx, rot_buf, rot_buf, pairs.data(), pairs.size(),
                           yb.data(), yb.size(), a_len, c_len, flags, out);
}

// ---------------------------------------------------------- block products
struct b2g_gemmlist_s {
    b2g_ctx ctx = nullptr;
    std::vector<Task> tasks[kTileConfigs];
    std::vector<Seg> segs;
    Task *d_tasks[kTileConfigs] = {};
    Seg *d_segs = nullptr;
    void *dev = nullptr;
    uint64_t d_len = 0, macs = 0, n_products = 0;
};

extern "C" int b2g_contract_wfn_plan(b2g_ctx ctx, int symmetry,
                                     const b2g_sparse_info *lmat, int l_is_wfn,
                                     const b2g_sparse_info *rmat,
                                     const b2g_sparse_info *outi, double scale,
                                     b2g_gemmlist *plan) {
    B2G_REQUIRE(ctx && lmat && rmat && outi && plan, B2G_ERR_ARG, "NULL arg");
    B2G_REQUIRE(symmetry == B2G_SZ || symmetry == B2G_SU2, B2G_ERR_ARG,
                "bad symmetry");
    const Symm s = symm_of(symmetry);
    BlockInfo L, R, O;
    int r;
    if ((r = to_block(*lmat, "lmat", L)) != B2G_OK ||
        (r = to_block(*rmat, "rmat", R)) != B2G_OK ||
        (r = to_block(*outi, "out", O)) != B2G_OK)
        return r;
    std::unique_ptr<b2g_gemmlist_s> g(new b2g_gemmlist_s());
    g->ctx = ctx;
    g->d_len = O.total();
    tile_env_refresh();
    for (int i = 0; i < O.n(); i++) {
        int il, ir;
        if (l_is_wfn) { // wfn = wfn x rot
            il = L.find_state(O.quanta[i]);
            ir = R.find_state(q_neg(s, q_get_ket(s, O.quanta[i])));
        } else { // wfn = rot x wfn
            il = L.find_state(q_get_bra(s, O.quanta[i], O.dq));
            ir = R.find_state(O.quanta[i]);
        }
        if (il < 0 || ir < 0)
            continue;
        const uint32_t M = O.nbra[i], N = O.nket[i], K = L.nket[il];
        B2G_REQUIRE(L.nbra[il] == M && R.nket[ir] == N && R.nbra[ir] == K,
                    B2G_ERR_ARG, "contraction block shapes differ");
        if (K == 0 || M == 0 || N == 0)
            continue;
        const uint32_t sb = (uint32_t)g->segs.size();
        Seg sg{};
        sg.a_off = L.shift[il];
        sg.b_off = R.shift[ir];
        sg.lda = L.nket[il];
        sg.ldb = R.nket[ir];
        sg.K = K;
        sg.flags = 0;
        sg.alpha = scale;
        g->segs.push_back(sg);
        add_tiles(g->tasks, O.shift[i], M, N, K, sb, sb + 1, 0);
        g->macs += (uint64_t)M * N * K;
        g->n_products++;
    }
    size_t bytes = g->segs.size() * sizeof(Seg);
    for (int c = 0; c < kTileConfigs; c++)
        bytes += g->tasks[c].size() * sizeof(Task);
    if (bytes) {
        B2G_HIP(hipSetDevice(ctx->device));
        B2G_HIP(hipMalloc(&g->dev, bytes));
        std::vector<char> host(bytes);
        size_t off = 0;
        for (int c = 0; c < kTileConfigs; c++) {
            const size_t nb = g->tasks[c].size() * sizeof(Task);
            if (nb)
                std::memcpy(host.data() + off, g->tasks[c].data(), nb);
            g->d_tasks[c] = (Task *)((char *)g->dev + off);
            off += nb;
        }
        std::memcpy(host.data() + off, g->segs.data(),
               g->segs.size() * sizeof(Seg));
        g->d_segs = (Seg *)((char *)g->dev + off);
        B2G_HIP(hipMemcpy(g->dev, host.data(), bytes, hipMemcpyHostToDevice));
    }
    *plan = g.release();
    return B2G_OK;
}

extern "C" int b2g_gemmlist_run(b2g_gemmlist g, const double *l,
                                const double *r, double *d) {
    B2G_REQUIRE(g && d && (g->segs.empty() || (l && r)), B2G_ERR_ARG,
                "NULL arg");
    B2G_HIP(hipSetDevice(g->ctx->device));
    hipStream_t st = g->ctx->stream;
    if (g->d_len)
        B2G_HIP(launch_fill(0.0, d, g->d_len, st));
    const int grid = g->ctx->num_cu * 8;
    for (int c = 0; c < kTileConfigs; c++) {
        const int nt = (int)g->tasks[c].size();
        if (nt == 0)
            continue;
        B2G_HIP((launch_seg_gemm<double, double, double, double, 6>(
            c, l, r, d, d, g->d_tasks[c], nt, g->d_segs, 0, 0, 0, 0, 1.0, 0, 1,
            grid, st)));
    }
    return B2G_OK;
}

extern "C" int b2g_gemmlist_stats(b2g_gemmlist g, uint64_t *n_products,
                                  uint64_t *macs) {
    B2G_REQUIRE(g, B2G_ERR_ARG, "NULL plan");
    if (n_products)
        *n_products = g->n_products;
    if (macs)
        *macs = g->macs;
    return B2G_OK;
}

extern "C" int b2g_gemmlist_destroy(b2g_gemmlist g) {
    if (g == nullptr)
        return B2G_OK;
    if (g->dev) {
        hipSetDevice(g->ctx->device);
        hipStreamSynchronize(g->ctx->stream);
        hipFree(g->dev);
    }
    delete g;
    return B2G_OK;
}
