// Effective-Hamiltonian layer: problem intake, plan building, diagonal and
// the Davidson solve, on top of the thin device layer.
//
// Mirrors EffectiveHamiltonian<S, FL, MPS> (src/dmrg/effective_hamiltonian.hpp):
//   ctor        -> b2g_effham_create: sub-labels, ConnectionInfo for the
//                  diagonal and the wavefunction, diagonal (:126-224)
//   precompute  -> pair recording + upload (:226-246)
//   operator()  -> b2g_effham_hpsi (:459-477)
//   eigs        -> b2g_effham_eigs (:480-568)
//   deallocate  -> b2g_effham_destroy (:1136-1162)
#include "common.hpp"
#include "launch.hpp"
#include "plan.hpp"
#include "replay.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <unordered_map>

using namespace b2g;

namespace b2g {
// Graph-captured Davidson: workspace and one instantiated iteration per
// precision, kept across eigs calls.  A graph is re-captured when the plan's
// scratch moved (scratch_gen) or the subspace size changed; the operator
// data it reads is updated in place and stays valid.
struct DavCache {
    size_t n = 0, nv = 0;
    double *B = nullptr, *S = nullptr, *q = nullptr, *t = nullptr;
    double *qn = nullptr, *sn = nullptr, *partial = nullptr;
    DavState *dst = nullptr, *hdr = nullptr; // hdr: pinned host
    // per precision: [1] product + step, [0] step alone
    hipGraphExec_t exec[3][2] = {};
    uint64_t gen[3] = {};
    void drop_graphs() {
        for (auto &ep : exec)
            for (auto &e : ep)
                if (e) {
                    hipGraphExecDestroy(e);
                    e = nullptr;
                }
    }
    void release() {
        drop_graphs();
        for (double **p : {&B, &S, &q, &t, &qn, &sn, &partial})
            if (*p) {
                hipFree(*p);
                *p = nullptr;
            }
        if (dst)
            hipFree(dst);
        if (hdr)
            hipHostFree(hdr);
        dst = hdr = nullptr;
        n = nv = 0;
    }
    ~DavCache() { release(); }
};
} // namespace b2g

struct b2g_effham_s {
    b2g_ctx ctx = nullptr;
    Problem prob;
    Connections wfn, dg;
    std::vector<PairRec> pairs;
    std::vector<DiagRec> drecs;
    b2g_buf lbuf = nullptr, rbuf = nullptr;
    b2g_plan plan = nullptr;
    double *diag = nullptr; // device, bra length
    uint64_t macs = 0;
    // Davidson driver: 1 graph-captured device iteration (one root),
    // 0 host-driven loop
    int dav_mode = 1;
    // delayed strings present (their 1 x 1 sub-operator values are folded
    // into the recorded factors, so the numbers cannot be replaced)
    bool has_delayed = false;
    // optional matvec hook (EffectiveKernel)
    b2g_kernel_fn kernel = nullptr;
    void *kernel_user = nullptr;
    b2g::DavCache dav;
};

namespace b2g {

static BlockInfo to_info(const b2g_sparse_info &s) {
    BlockInfo b;
    b.dq = s.delta_quantum;
    b.is_fermion = s.is_fermion != 0;
    b.quanta.assign(s.quanta, s.quanta + s.n);
    b.nbra.assign(s.n_states_bra, s.n_states_bra + s.n);
    b.nket.assign(s.n_states_ket, s.n_states_ket + s.n);
    b.shift.assign(s.block_shifts, s.block_shifts + s.n);
    return b;
}

static int check_info(const b2g_sparse_info &s, const char *what) {
    B2G_REQUIRE(s.n >= 0, B2G_ERR_ARG, std::string(what) + ": negative n");
    B2G_REQUIRE(s.n == 0 || (s.quanta && s.n_states_bra && s.n_states_ket &&
                             s.block_shifts),
                B2G_ERR_ARG, std::string(what) + ": NULL arrays");
    for (int i = 1; i < s.n; i++)
        B2G_REQUIRE(s.quanta[i - 1] < s.quanta[i], B2G_ERR_ARG,
                    std::string(what) + ": quanta not strictly ascending");
    return B2G_OK;
}

static std::vector<b2g_block> blocks_of(const BlockInfo &b) {
    std::vector<b2g_block> r(b.n());
    for (int i = 0; i < b.n(); i++) {
        r[i].off = b.shift[i];
        r[i].m = b.nbra[i];
        r[i].n = b.nket[i];
    }
    return r;
}

// Symmetric eigen-decomposition of an m x m matrix (cyclic Jacobi,
// FP64).  On return w holds ascending eigenvalues and row j of v the
// eigenvector of w[j] (the layout GMatrixFunctions::eigs leaves in its
// input, src/core/matrix_functions.hpp:1672-1688).
void jacobi_eigs(int m, std::vector<double> a, std::vector<double> &w,
                 std::vector<double> &v) {
    std::vector<double> z(m * m, 0.0);
    for (int i = 0; i < m; i++)
        z[i * m + i] = 1.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) {
                tot += a[i * m + j] * a[i * m + j];
                if (i != j)
                    off += a[i * m + j] * a[i * m + j];
            }
        if (off <= 1E-30 * tot || off == 0.0)
            break;
        for (int p = 0; p < m - 1; p++)
            for (int q = p + 1; q < m; q++) {
                const double apq = a[p * m + q];
                if (std::abs(apq) < 1E-300)
                    continue;
                const double app = a[p * m + p], aqq = a[q * m + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t =
                    (theta >= 0 ? 1.0 : -1.0) /
                    (std::abs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < m; k++) {
                    const double akp = a[k * m + p], akq = a[k * m + q];
                    a[k * m + p] = c * akp - s * akq;
                    a[k * m + q] = s * akp + c * akq;
                }
                for (int k = 0; k < m; k++) {
                    const double apk = a[p * m + k], aqk = a[q * m + k];
                    a[p * m + k] = c * apk - s * aqk;
                    a[q * m + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < m; k++) {
                    const double zkp = z[k * m + p], zkq = z[k * m + q];
                    z[k * m + p] = c * zkp - s * zkq;
                    z[k * m + q] = s * zkp + c * zkq;
                }
            }
    }
    std::vector<int> idx(m);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&a, m](int x, int y) {
        return a[x * m + x] < a[y * m + y];
    });
    w.resize(m);
    v.assign(m * m, 0.0);
    for (int j = 0; j < m; j++) {
        w[j] = a[idx[j] * m + idx[j]];
        for (int i = 0; i < m; i++)
            v[j * m + i] = z[i * m + idx[j]];
    }
}

} // namespace b2g

namespace b2g {

// diagonal of H from the current operator buffers
static int update_diag(b2g_effham h) {
    std::vector<b2g_diag_term> dt(h->drecs.size());
    for (size_t i = 0; i < dt.size(); i++) {
        const DiagRec &q = h->drecs[i];
        dt[i] = b2g_diag_term{q.a_off, q.b_off, q.d_off, q.a_n,
                              q.b_n,   q.r0,    q.c0,    q.alpha};
    }
    std::vector<b2g_block> db = blocks_of(h->prob.ket);
    return b2g_diag(h->ctx, h->lbuf, h->rbuf, dt.data(), dt.size(), db.data(),
                    db.size(), h->diag, h->prob.ket.total());
}

} // namespace b2g

extern "C" int b2g_effham_create(b2g_ctx ctx, const b2g_problem *pb,
                                 unsigned flags, b2g_effham *out) {
    B2G_REQUIRE((ctx || (flags & B2G_EFFHAM_HOST_ONLY)) && pb && out,
                B2G_ERR_ARG, "NULL arg");
    B2G_REQUIRE(pb->symmetry == B2G_SZ || pb->symmetry == B2G_SU2,
                B2G_ERR_ARG, "unknown symmetry");
    int r;
    if ((r = check_info(pb->ket, "ket")) != B2G_OK ||
        (r = check_info(pb->bra, "bra")) != B2G_OK)
        return r;
    B2G_REQUIRE(pb->n_left_infos >= 0 && pb->n_right_infos >= 0 &&
                    pb->n_left_ops >= 0 && pb->n_right_ops >= 0 &&
                    pb->n_terms >= 0,
                B2G_ERR_ARG, "negative counts");
    SetupTimer tm("effham");
    std::unique_ptr<b2g_effham_s> h(new b2g_effham_s());
    h->ctx = ctx;
    Problem &p = h->prob;
    p.symm = pb->symmetry == B2G_SZ ? Symm::SZ : Symm::SU2;
    p.opdq = pb->opdq;
    p.hop_left_vacuum = pb->hop_left_vacuum;
    p.ket = to_info(pb->ket);
    p.bra = to_info(pb->bra);
    for (int i = 0; i < pb->n_left_infos; i++) {
        if ((r = check_info(pb->left_infos[i], "left info")) != B2G_OK)
            return r;
        B2G_REQUIRE(i == 0 || pb->left_info_dq[i - 1] < pb->left_info_dq[i],
                    B2G_ERR_ARG, "left infos not sorted by delta quantum");
        p.linfos.push_back(to_info(pb->left_infos[i]));
        p.linfos.back().dq = pb->left_info_dq[i];
    }
    for (int i = 0; i < pb->n_right_infos; i++) {
        if ((r = check_info(pb->right_infos[i], "right info")) != B2G_OK)
            return r;
        B2G_REQUIRE(i == 0 || pb->right_info_dq[i - 1] < pb->right_info_dq[i],
                    B2G_ERR_ARG, "right infos not sorted by delta quantum");
        p.rinfos.push_back(to_info(pb->right_infos[i]));
        p.rinfos.back().dq = pb->right_info_dq[i];
    }
    auto take_ops = [](const b2g_operator *ops, int n,
                       const std::vector<BlockInfo> &infos, size_t len,
                       std::vector<Operator> &dst, const char *what) -> int {
        for (int i = 0; i < n; i++) {
            const b2g_operator &o = ops[i];
            B2G_REQUIRE(o.info >= 0 && o.info < (int)infos.size(), B2G_ERR_ARG,
                        std::string(what) + " operator: bad info index");
            B2G_REQUIRE(infos[o.info].dq == o.delta_quantum, B2G_ERR_ARG,
                        std::string(what) +
                            " operator: delta quantum differs from its info");
            B2G_REQUIRE((o.flags & 1) ||
                            o.data_off + infos[o.info].total() <= len,
                        B2G_ERR_ARG,
                        std::string(what) + " operator: data out of range");
            dst.push_back(Operator{o.delta_quantum, o.factor, o.info,
                                   o.data_off});
        }
        return B2G_OK;
    };
    if ((r = take_ops(pb->left_ops, pb->n_left_ops, p.linfos, pb->left_len,
                      p.lops, "left")) != B2G_OK ||
        (r = take_ops(pb->right_ops, pb->n_right_ops, p.rinfos, pb->right_len,
                      p.rops, "right")) != B2G_OK)
        return r;
    // delayed tensor: sub-operator sets and tensor-product connection lists
    B2G_REQUIRE(pb->delayed_side >= 0 && pb->delayed_side <= 2, B2G_ERR_ARG,
                "delayed_side must be 0, 1 or 2");
    p.delayed_side = pb->delayed_side;
    if (p.delayed_side != 0) {
        B2G_REQUIRE(pb->n_dl_infos >= 0 && pb->n_dr_infos >= 0 &&
                        pb->n_dl_ops >= 0 && pb->n_dr_ops >= 0 && pb->tp,
                    B2G_ERR_ARG, "delayed tensor: bad counts or NULL tp");
        for (int i = 0; i < pb->n_dl_infos; i++) {
            if ((r = check_info(pb->dl_infos[i], "delayed left info")) != B2G_OK)
                return r;
            p.dlinfos.push_back(to_info(pb->dl_infos[i]));
        }
        for (int i = 0; i < pb->n_dr_infos; i++) {
            if ((r = check_info(pb->dr_infos[i], "delayed right info")) !=
                B2G_OK)
                return r;
            p.drinfos.push_back(to_info(pb->dr_infos[i]));
        }
        const size_t dlen =
            p.delayed_side == 1 ? pb->left_len : pb->right_len;
        if ((r = take_ops(pb->dl_ops, pb->n_dl_ops, p.dlinfos, dlen, p.dlops,
                          "delayed left")) != B2G_OK ||
            (r = take_ops(pb->dr_ops, pb->n_dr_ops, p.drinfos, dlen, p.drops,
                          "delayed right")) != B2G_OK)
            return r;
        const std::vector<BlockInfo> &side =
            p.delayed_side == 1 ? p.linfos : p.rinfos;
        p.tp.resize(side.size());
        for (size_t i = 0; i < side.size(); i++) {
            const b2g_tp_conn &c = pb->tp[i];
            Connections &d = p.tp[i];
            if (c.nc <= 0)
                continue;
            B2G_REQUIRE(c.n[4] > 0 && c.quanta && c.idx && c.ia && c.ib &&
                            c.ic && c.stride && c.factor,
                        B2G_ERR_ARG, "tensor-product list: NULL arrays");
            for (int k = 0; k < 5; k++)
                d.n[k] = c.n[k];
            d.quanta.assign(c.quanta, c.quanta + c.n[4]);
            d.idx.assign(c.idx, c.idx + c.n[4]);
            d.ia.assign(c.ia, c.ia + c.nc);
            d.ib.assign(c.ib, c.ib + c.nc);
            d.ic.assign(c.ic, c.ic + c.nc);
            d.stride.assign(c.stride, c.stride + c.nc);
            d.factor.assign(c.factor, c.factor + c.nc);
            for (int k = 0; k < c.n[4]; k++)
                B2G_REQUIRE(d.idx[k] <= (uint32_t)c.nc, B2G_ERR_ARG,
                            "tensor-product list: class index out of range");
            for (int k = 0; k < c.nc; k++)
                B2G_REQUIRE(d.ic[k] < (uint32_t)side[i].n(), B2G_ERR_ARG,
                            "tensor-product list: fused block out of range");
        }
    }
    for (int i = 0; i < pb->n_terms; i++) {
        const b2g_term &t = pb->terms[i];
        B2G_REQUIRE(t.a >= 0 && t.a < pb->n_left_ops && t.b >= 0 &&
                        t.b < pb->n_right_ops && t.conj < 4 && t.kind < 2,
                    B2G_ERR_ARG, "term references a missing operator");
        Term tm{t.a, t.b, (uint8_t)t.conj, t.factor};
        if (t.kind == 1) {
            B2G_REQUIRE(p.delayed_side != 0 && t.da >= 0 &&
                            t.da < (int)p.dlops.size() && t.db >= 0 &&
                            t.db < (int)p.drops.size() && t.dconj < 4,
                        B2G_ERR_ARG,
                        "delayed term references a missing sub-operator");
            tm.three = true;
            tm.da = t.da, tm.db = t.db, tm.dconj = (uint8_t)t.dconj;
        }
        p.terms.push_back(tm);
    }
    tm.mark("problem");
    // connection lists (EffectiveHamiltonian ctor)
    SU2Coupling cg;
    std::vector<SubLabel> subl = uniq_sub_labels(p);
    std::vector<InfoRef> ai, bi;
    for (auto &x : p.linfos)
        ai.emplace_back(x.dq, &x);
    for (auto &x : p.rinfos)
        bi.emplace_back(x.dq, &x);
    build_diag_connections(h->dg, p.symm, cg, p.ket.dq, p.opdq, subl, ai, bi,
                           p.ket);
    build_wfn_connections(h->wfn, p.symm, cg, p.ket.dq, p.bra.dq, p.opdq, subl,
                          ai, bi, p.ket, p.bra);
    tm.mark("connection lists");
    std::string err;
    h->pairs = record_pairs(p, h->wfn, err);
    B2G_REQUIRE(err.empty(), B2G_ERR_ARG, err);
    h->drecs = record_diag(p, h->dg, err);
    B2G_REQUIRE(err.empty(), B2G_ERR_ARG, err);
    // fold the 1 x 1 delayed sub-operator values into the factors
    std::vector<int64_t> soff;
    for (auto &q : h->pairs)
        if (q.scal_off >= 0)
            soff.push_back(q.scal_off);
    for (auto &q : h->drecs)
        if (q.scal_off >= 0)
            soff.push_back(q.scal_off);
    if (!soff.empty()) {
        const double *src =
            p.delayed_side == 1 ? pb->left_data : pb->right_data;
        std::sort(soff.begin(), soff.end());
        soff.erase(std::unique(soff.begin(), soff.end()), soff.end());
        std::unordered_map<int64_t, double> sval;
        if (src != nullptr && (flags & B2G_EFFHAM_DEVICE_DATA)) {
            B2G_REQUIRE(ctx, B2G_ERR_ARG, "device data needs a context");
            std::vector<double> v(soff.size());
            for (size_t k = 0; k < soff.size(); k++)
                B2G_HIP(hipMemcpy(&v[k], src + soff[k], sizeof(double),
                                  hipMemcpyDeviceToHost));
            for (size_t k = 0; k < soff.size(); k++)
                sval[soff[k]] = v[k];
        } else if (src != nullptr) {
            for (int64_t o : soff)
                sval[o] = src[o];
        }
        h->has_delayed = true;
        if (!sval.empty()) {
            for (auto &q : h->pairs)
                if (q.scal_off >= 0)
                    q.alpha *= sval[q.scal_off];
            for (auto &q : h->drecs)
                if (q.scal_off >= 0)
                    q.alpha *= sval[q.scal_off];
        }
    }
    h->macs = 0;
    for (auto &q : h->pairs)
        h->macs += pair_macs(q);
    tm.mark("pair and diagonal records");
    if (flags & B2G_EFFHAM_HOST_ONLY) {
        // plan only: connection lists and pairs, no device state
        *out = h.release();
        return B2G_OK;
    }
    // upload operators and plan
    if (flags & B2G_EFFHAM_DEVICE_DATA) {
        if ((r = b2g_buf_create(ctx, pb->left_len, B2G_F64, &h->lbuf)) !=
            B2G_OK)
            return r;
        if ((r = b2g_buf_create(ctx, pb->right_len, B2G_F64, &h->rbuf)) !=
            B2G_OK) {
            b2g_buf_destroy(h->lbuf);
            return r;
        }
        if (hipMemcpyAsync(h->lbuf->ptr, pb->left_data,
                           pb->left_len * sizeof(double),
                           hipMemcpyDeviceToDevice, ctx->stream) !=
                hipSuccess ||
            hipMemcpyAsync(h->rbuf->ptr, pb->right_data,
                           pb->right_len * sizeof(double),
                           hipMemcpyDeviceToDevice, ctx->stream) !=
                hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            b2g_buf_destroy(h->lbuf);
            b2g_buf_destroy(h->rbuf);
            set_error("copy of device operator data failed");
            return B2G_ERR_DEVICE;
        }
    } else {
        if ((r = b2g_upload_blocks(ctx, pb->left_data, pb->left_len, B2G_F64,
                                   &h->lbuf)) != B2G_OK)
            return r;
        if ((r = b2g_upload_blocks(ctx, pb->right_data, pb->right_len,
                                   B2G_F64, &h->rbuf)) != B2G_OK) {
            b2g_buf_destroy(h->lbuf);
            return r;
        }
    }
    tm.mark("operator upload");
    std::vector<b2g_pair> dp(h->pairs.size());
    for (size_t i = 0; i < dp.size(); i++) {
        const PairRec &q = h->pairs[i];
        dp[i] = b2g_pair{q.a_off, q.b_off, q.c_off, q.v_off, q.c_m, q.c_n,
                         q.v_m, q.v_n, q.conj, q.kind, q.c_ld, q.v_ld,
                         q.alpha};
    }
    std::vector<b2g_block> yb = blocks_of(p.bra);
    r = b2g_upload_plan(ctx, h->lbuf, h->rbuf, dp.data(), dp.size(), yb.data(),
                        yb.size(), p.ket.total(), p.bra.total(), flags,
                        &h->plan);
    tm.mark("plan upload");
    if (r == B2G_OK) {
        const size_t dl = std::max<uint64_t>(p.ket.total(), 1);
        if (hipMalloc(&h->diag, dl * sizeof(double)) != hipSuccess) {
            set_error("diagonal allocation failed");
            r = B2G_ERR_MEMORY;
        } else
            r = update_diag(h.get());
    }
    tm.mark("diagonal");
    if (r != B2G_OK) {
        b2g_effham_destroy(h.release());
        return r;
    }
    *out = h.release();
    return B2G_OK;
}

extern "C" int b2g_effham_destroy(b2g_effham h) {
    if (h == nullptr)
        return B2G_OK;
    if (h->ctx)
        hipStreamSynchronize(h->ctx->stream);
    h->dav.release();
    b2g_plan_destroy(h->plan);
    b2g_buf_destroy(h->lbuf);
    b2g_buf_destroy(h->rbuf);
    if (h->diag)
        hipFree(h->diag);
    delete h;
    return B2G_OK;
}

extern "C" int b2g_effham_diag_stats(b2g_effham h, uint64_t *n_terms,
                                     uint64_t *window_elems,
                                     uint64_t *read_elems) {
    B2G_REQUIRE(h, B2G_ERR_ARG, "NULL effham");
    uint64_t w = 0, r = 0;
    for (const DiagRec &q : h->drecs) {
        w += (uint64_t)q.a_n * q.b_n;
        r += (uint64_t)q.a_n + q.b_n;
    }
    if (n_terms)
        *n_terms = h->drecs.size();
    if (window_elems)
        *window_elems = w;
    if (read_elems)
        *read_elems = r;
    return B2G_OK;
}

extern "C" int b2g_effham_info(b2g_effham h, uint64_t *ket_len,
                               uint64_t *bra_len, uint64_t *n_pairs,
                               uint64_t *macs) {
    B2G_REQUIRE(h, B2G_ERR_ARG, "NULL effham");
    if (ket_len)
        *ket_len = h->prob.ket.total();
    if (bra_len)
        *bra_len = h->prob.bra.total();
    if (n_pairs)
        *n_pairs = h->pairs.size();
    if (macs)
        *macs = h->macs;
    return B2G_OK;
}

extern "C" int b2g_effham_connections(b2g_effham h, int which,
                                      uint64_t *n_conn, uint32_t *ia,
                                      uint32_t *ib, uint32_t *ic,
                                      uint64_t *stride, double *factor) {
    B2G_REQUIRE(h && n_conn && (which == 0 || which == 1), B2G_ERR_ARG,
                "bad arg");
    const Connections &c = which == 0 ? h->wfn : h->dg;
    *n_conn = c.ia.size();
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

extern "C" int b2g_effham_pairs(b2g_effham h, uint64_t *n_pairs,
                                b2g_pair *pairs) {
    B2G_REQUIRE(h && n_pairs, B2G_ERR_ARG, "NULL arg");
    *n_pairs = h->pairs.size();
    if (pairs)
        for (size_t i = 0; i < h->pairs.size(); i++) {
            const PairRec &q = h->pairs[i];
            pairs[i] = b2g_pair{q.a_off, q.b_off, q.c_off, q.v_off, q.c_m,
                                q.c_n, q.v_m, q.v_n, q.conj, q.kind,
                                q.c_ld, q.v_ld, q.alpha};
        }
    return B2G_OK;
}

extern "C" b2g_plan b2g_effham_plan(b2g_effham h) {
    return h ? h->plan : nullptr;
}

extern "C" int b2g_effham_hpsi(b2g_effham h, const double *x, double *y,
                               int n_rhs, double scale, int accumulate,
                               int precision) {
    B2G_REQUIRE(h, B2G_ERR_ARG, "NULL effham");
    B2G_REQUIRE(h->plan, B2G_ERR_STATE, "effham was built host-only");
    return b2g_hpsi(h->ctx, h->plan, x, y, n_rhs, scale, accumulate,
                    precision);
}

extern "C" int b2g_effham_diag(b2g_effham h, double *d) {
    B2G_REQUIRE(h && d, B2G_ERR_ARG, "NULL arg");
    B2G_REQUIRE(h->diag, B2G_ERR_STATE, "effham was built host-only");
    B2G_HIP(hipMemcpyAsync(d, h->diag, h->prob.ket.total() * sizeof(double),
                           hipMemcpyDeviceToDevice, h->ctx->stream));
    return B2G_OK;
}

extern "C" int b2g_effham_set_data(b2g_effham h, const double *left,
                                   const double *right) {
    B2G_REQUIRE(h && left && right, B2G_ERR_ARG, "NULL arg");
    B2G_REQUIRE(h->plan, B2G_ERR_STATE, "effham was built host-only");
    B2G_REQUIRE(!h->has_delayed, B2G_ERR_STATE,
                "operator data of a problem with delayed strings are part "
                "of its plan; create a new effective Hamiltonian");
    B2G_HIP(hipSetDevice(h->ctx->device));
    if (h->lbuf->n)
        B2G_HIP(hipMemcpyAsync(h->lbuf->ptr, left, h->lbuf->n * sizeof(double),
                               hipMemcpyDeviceToDevice, h->ctx->stream));
    if (h->rbuf->n)
        B2G_HIP(hipMemcpyAsync(h->rbuf->ptr, right,
                               h->rbuf->n * sizeof(double),
                               hipMemcpyDeviceToDevice, h->ctx->stream));
    int r = b2g_plan_refresh(h->plan);
    if (r == B2G_OK)
        r = update_diag(h);
    return r;
}

extern "C" int b2g_effham_data(b2g_effham h, double **left, size_t *left_len,
                               double **right, size_t *right_len) {
    B2G_REQUIRE(h && left && left_len && right && right_len, B2G_ERR_ARG,
                "NULL arg");
    B2G_REQUIRE(h->plan, B2G_ERR_STATE, "effham was built host-only");
    *left = (double *)h->lbuf->ptr;
    *left_len = h->lbuf->n;
    *right = (double *)h->rbuf->ptr;
    *right_len = h->rbuf->n;
    return B2G_OK;
}

extern "C" int b2g_effham_set_kernel(b2g_effham h, b2g_kernel_fn kernel,
                                     void *user) {
    B2G_REQUIRE(h, B2G_ERR_ARG, "NULL effham");
    h->kernel = kernel;
    h->kernel_user = user;
    return B2G_OK;
}

extern "C" int b2g_effham_set_davidson_mode(b2g_effham h, int mode) {
    B2G_REQUIRE(h && (mode == 0 || mode == 1), B2G_ERR_ARG, "bad mode");
    h->dav_mode = mode;
    return B2G_OK;
}

extern "C" int b2g_effham_grad_ops(b2g_effham h, const double *v,
                                   const double *c, int n_rhs, double *gl,
                                   double *gr, int accumulate) {
    B2G_REQUIRE(h, B2G_ERR_ARG, "NULL effham");
    B2G_REQUIRE(h->plan, B2G_ERR_STATE, "effham was built host-only");
    B2G_REQUIRE(!h->has_delayed, B2G_ERR_STATE,
                "gradients of problems with delayed strings are not formed "
                "(their 1 x 1 sub-operator values are plan constants)");
    return b2g_grad_ops(h->ctx, h->plan, v, c, n_rhs, gl, gr, accumulate);
}

// --------------------------------------------------------------- Davidson
namespace b2g {

struct DavWork {
    double *B = nullptr, *S = nullptr, *q = nullptr, *t = nullptr;
    double *O = nullptr;    // orthogonal states
    double *scal = nullptr; // device scalars (>= 64*64 + 16)
    ~DavWork() {
        if (O)
            hipFree(O);
        if (B)
            hipFree(B);
        if (S)
            hipFree(S);
        if (q)
            hipFree(q);
        if (t)
            hipFree(t);
        if (scal)
            hipFree(scal);
    }
};

static int dev_dot(b2g_ctx ctx, const double *x, const double *y, size_t n,
                   double &out) {
    return b2g_vec_dot(ctx, x, y, n, &out);
}

struct MatvecData {
    b2g_effham h;
    int precision;
};

static int default_matvec(void *d, const double *a, double *b, double scale) {
    const MatvecData *m = (const MatvecData *)d;
    return b2g_hpsi(m->h->ctx, m->h->plan, a, b, 1, scale, 0, m->precision);
}

// b = H a, through the user hook when one is set
static int davidson_matvec(b2g_effham h, int precision, const double *a,
                           double *b) {
    if (h->kernel == nullptr)
        return b2g_hpsi(h->ctx, h->plan, a, b, 1, 1.0, 0, precision);
    B2G_HIP(hipStreamSynchronize(h->ctx->stream));
    MatvecData md{h, precision};
    const int r = h->kernel(h->kernel_user, 1.0, default_matvec, &md, a, b,
                            h->prob.ket.total());
    B2G_REQUIRE(r == 0, B2G_ERR_STATE, "the matvec hook returned an error");
    return B2G_OK;
}

// Davidson options shared by the host loop and the graph: the arguments of
// EffectiveHamiltonian::eigs (src/dmrg/effective_hamiltonian.hpp:480-490).
struct DavOpts {
    double conv = 5E-6, rel = 0.0, shift = 0.0;
    int max_iter = 5000, soft_max_iter = -1, def_min = 2, def_max = 50;
    int type = B2G_DAV_NORMAL;
    int n_ortho = 0;
    const double *const *ortho = nullptr; // device vectors
    const double *proj_weights = nullptr; // host, n_ortho values or nullptr
};

// Order of the Ritz values a Davidson type targets
// (iterative_matrix_functions.hpp:1031-1061): ascending for Normal, else
// closest to / at or below / above the shift first.  Stable.
static void ritz_order(int type, double shift, const std::vector<double> &ld,
                       int m, std::vector<int> &idx) {
    idx.resize(m);
    std::iota(idx.begin(), idx.end(), 0);
    auto before = [type, shift, &ld](int i, int j) {
        const double a = ld[i], b = ld[j];
        if (type & B2G_DAV_CLOSE_TO)
            return std::abs(a - shift) < std::abs(b - shift);
        if (type & B2G_DAV_LESS_THAN) {
            if ((shift >= a) != (shift >= b))
                return shift >= a;
            if (shift >= a)
                return shift - a < shift - b;
            return a - shift > b - shift;
        }
        if (type & B2G_DAV_GREATER_THAN) {
            if ((shift > a) != (shift > b))
                return shift > b;
            if (shift > a)
                return shift - a > shift - b;
            return a - shift < b - shift;
        }
        return false;
    };
    if (type & (B2G_DAV_CLOSE_TO | B2G_DAV_LESS_THAN | B2G_DAV_GREATER_THAN))
        std::stable_sort(idx.begin(), idx.end(), before);
}

// One Davidson solve with the vectors on the device and the m x m Ritz
// problem on the host (IterativeMatrixFunctions::davidson,
// src/core/iterative_matrix_functions.hpp:866-1173): the k roots' products
// of one iteration run as one multi-RHS launch; DavidsonTypes Normal,
// GreaterThan, LessThan and CloseTo with the Olsen (:93-108), Davidson
// (:66-72) or no preconditioner; orthogonal states projected out or, with
// projection weights, added to H as w|v><v|; rel_conv_thrd and
// soft_max_iter.
static int davidson(b2g_effham h, double *x, int k, const DavOpts &o,
                    int precision, std::vector<double> &eigvals, int &ndav,
                    uint64_t &nmv, bool &soft_stop) {
    b2g_ctx ctx = h->ctx;
    hipStream_t st = ctx->stream;
    const size_t n = h->prob.ket.total();
    B2G_REQUIRE(h->prob.bra.total() == n, B2G_ERR_ARG,
                "Davidson needs bra and ket of equal size");
    B2G_REQUIRE(k >= 1 && k <= 32, B2G_ERR_ARG, "bad number of roots");
    int def_min = o.def_min, def_max = o.def_max;
    if (def_min < k)
        def_min = k;
    if (def_max < k + k / 2)
        def_max = k + k / 2;
    B2G_REQUIRE(def_max <= 64 && def_min < def_max, B2G_ERR_ARG,
                "deflation sizes must satisfy k <= def_min < def_max <= 64");
    int nor = o.n_ortho, nwg = 0;
    B2G_REQUIRE(nor >= 0 && nor <= 64 && (nor == 0 || o.ortho), B2G_ERR_ARG,
                "bad orthogonal states");
    for (int j = 0; j < nor; j++)
        B2G_REQUIRE(o.ortho[j], B2G_ERR_ARG, "NULL orthogonal state");
    if (nor > 0 && o.proj_weights && !(o.type & B2G_DAV_ELEMENT_PROJ))
        nwg = nor, nor = 0;
    const int nv = std::max(nor, nwg);
    const double rel2 = o.rel * o.rel;
    DavWork w;
    B2G_HIP(hipMalloc(&w.B, (size_t)def_max * n * sizeof(double)));
    B2G_HIP(hipMalloc(&w.S, (size_t)def_max * n * sizeof(double)));
    B2G_HIP(hipMalloc(&w.q, n * sizeof(double)));
    B2G_HIP(hipMalloc(&w.t, n * sizeof(double)));
    B2G_HIP(hipMalloc(&w.scal, (64 * 64 + 64) * sizeof(double)));
    if (nv)
        B2G_HIP(hipMalloc(&w.O, (size_t)nv * n * sizeof(double)));
    double *Bv = w.B, *Sv = w.S;
    auto bv = [&](int i) { return Bv + (size_t)i * n; };
    auto sv = [&](int i) { return Sv + (size_t)i * n; };
    auto ov = [&](int i) { return w.O + (size_t)i * n; };
    int r;
    // the orthogonal states, made orthogonal among themselves (:912-919)
    std::vector<double> orn(nv, 0.0);
    for (int i = 0; i < nv; i++)
        B2G_HIP(launch_copy(o.ortho[i], ov(i), n, st));
    for (int i = 0; i < nor; i++) {
        for (int j = 0; j < i; j++)
            if (std::abs(orn[j]) > 1E-14) {
                double d;
                if ((r = dev_dot(ctx, ov(j), ov(i), n, d)))
                    return r;
                B2G_HIP(launch_axpy(-d / orn[j], ov(j), ov(i), n, st));
            }
        if ((r = dev_dot(ctx, ov(i), ov(i), n, orn[i])))
            return r;
    }
    // q -= sum_j <o_j|q> / |o_j|^2 o_j
    auto project = [&](double *v) -> int {
        for (int j = 0; j < nor; j++)
            if (std::abs(orn[j]) > 1E-14) {
                double d;
                int rr = dev_dot(ctx, ov(j), v, n, d);
                if (rr)
                    return rr;
                B2G_HIP(launch_axpy(-d / orn[j], ov(j), v, n, st));
            }
        return B2G_OK;
    };
    for (int i = 0; i < k; i++)
        B2G_HIP(launch_copy(x + (size_t)i * n, bv(i), n, st));
    int m = k;
    for (int i = 0; i < k; i++) {
        for (int j = 0; j < i; j++) {
            double d;
            if ((r = dev_dot(ctx, bv(j), bv(i), n, d)))
                return r;
            B2G_HIP(launch_axpy(-d, bv(j), bv(i), n, st));
        }
        double nrm2;
        if ((r = dev_dot(ctx, bv(i), bv(i), n, nrm2)))
            return r;
        if (std::abs(nrm2) < 1E-14 && i > 0) {
            m = i;
            break;
        }
        B2G_REQUIRE(std::abs(nrm2) >= 1E-14, B2G_ERR_ARG,
                    "Davidson initial guess has zero norm");
        B2G_HIP(launch_scale(1.0 / std::sqrt(nrm2), bv(i), n, st));
    }
    for (int i = 0; i < m; i++) {
        if ((r = project(bv(i))))
            return r;
        double nrm2;
        if ((r = dev_dot(ctx, bv(i), bv(i), n, nrm2)))
            return r;
        B2G_REQUIRE(std::abs(nrm2) >= 1E-14, B2G_ERR_ARG,
                    "cannot generate Davidson initial guess " +
                        std::to_string(i) +
                        " orthogonal to the given states");
        B2G_HIP(launch_scale(1.0 / std::sqrt(nrm2), bv(i), n, st));
    }
    eigvals.assign(k, 0.0);
    std::vector<double> ld, V, G(64 * 64), A;
    std::vector<int> idx(def_max);
    std::iota(idx.begin(), idx.end(), 0);
    int ck = 0, msig = 0, xiter = 0;
    double qq = 0.0;
    nmv = 0;
    soft_stop = false;
    bool converged = false;
    const bool ordered = (o.type & (B2G_DAV_CLOSE_TO | B2G_DAV_LESS_THAN |
                                    B2G_DAV_GREATER_THAN)) != 0;
    while (xiter < o.max_iter &&
           (o.soft_max_iter == -1 || xiter < o.soft_max_iter)) {
        xiter++;
        if (msig < m) {
            // the products of all new vectors: one multi-RHS launch (one
            // call per vector through a user hook)
            if (h->kernel == nullptr) {
                if ((r = b2g_hpsi(ctx, h->plan, bv(msig), sv(msig), m - msig,
                                  1.0, 0, precision)))
                    return r;
            } else
                for (int i = msig; i < m; i++)
                    if ((r = davidson_matvec(h, precision, bv(i), sv(i))))
                        return r;
            for (int i = msig; i < m; i++)
                for (int j = 0; j < nwg; j++) {
                    double d;
                    if ((r = dev_dot(ctx, ov(j), bv(i), n, d)))
                        return r;
                    B2G_HIP(launch_axpy(d * o.proj_weights[j], ov(j), sv(i),
                                        n, st));
                }
            nmv += m - msig;
            msig = m;
        }
        // subspace matrix alpha(i, j) = <b_i | sigma_j>, lower triangle
        B2G_HIP(launch_gram(Bv, m, Sv, m, n, n, ctx->red, w.scal, st));
        B2G_HIP(hipMemcpyAsync(G.data(), w.scal, (size_t)m * m * sizeof(double),
                               hipMemcpyDeviceToHost, st));
        B2G_HIP(hipStreamSynchronize(st));
        A.assign((size_t)m * m, 0.0);
        for (int i = 0; i < m; i++)
            for (int j = 0; j <= i; j++)
                A[i * m + j] = A[j * m + i] = G[i * m + j];
        jacobi_eigs(m, A, ld, V);
        // rotate b and sigma into the Ritz basis
        B2G_HIP(hipMemcpyAsync(w.scal, V.data(), (size_t)m * m * sizeof(double),
                               hipMemcpyHostToDevice, st));
        B2G_HIP(launch_rotate(Sv, n, n, m, w.scal, st));
        B2G_HIP(launch_rotate(Bv, n, n, m, w.scal, st));
        ritz_order(o.type, o.shift, ld, m, idx);
        // re-check roots already converged
        for (int i = 0; i < ck; i++) {
            const int ii = idx[i];
            B2G_HIP(launch_copy(sv(ii), w.q, n, st));
            B2G_HIP(launch_axpy(-ld[ii], bv(ii), w.q, n, st));
            double d;
            if ((r = dev_dot(ctx, w.q, w.q, n, d)))
                return r;
            if (std::abs(d) >= o.conv + ld[ii] * ld[ii] * rel2) {
                ck = i;
                break;
            }
        }
        const int ick = idx[ck];
        B2G_HIP(launch_copy(sv(ick), w.q, n, st));
        B2G_HIP(launch_axpy(-ld[ick], bv(ick), w.q, n, st));
        if ((r = project(w.q)))
            return r;
        if ((r = dev_dot(ctx, w.q, w.q, n, qq)))
            return r;
        if (o.type & B2G_DAV_DAVIDSON_PRECOND)
            B2G_HIP(launch_dav_precond(w.q, h->diag, ld[ick], n, st));
        else if (!(o.type & B2G_DAV_NO_PRECOND))
            B2G_HIP(launch_olsen(w.q, bv(ick), h->diag, ld[ick], w.t, n,
                                 ctx->red, w.scal + 64 * 64, st));
        for (int i = 0; i <= ck; i++)
            eigvals[i] = ld[idx[i]];
        for (int i = ck + 1; i < k; i++)
            eigvals[i] = 0.0;
        if (std::abs(qq) < o.conv + eigvals[ck] * eigvals[ck] * rel2 &&
            m >= k) {
            ck++;
            if (ck == k) {
                converged = true;
                break;
            }
        } else {
            bool deflate = false;
            if (m >= def_max) {
                m = msig = def_min;
                deflate = ordered;
            }
            if (deflate) {
                // keep the targeted Ritz pairs first (:1105-1133): vector j
                // of B and S becomes the one of idx[j], through a buffer
                double *T = nullptr;
                B2G_HIP(hipMalloc(&T, (size_t)m * n * sizeof(double)));
                std::unique_ptr<double, int (*)(void *)> tg(
                    T, [](void *p) { return (int)hipFree(p); });
                for (double *X : {Bv, Sv}) {
                    for (int j = 0; j < m; j++)
                        B2G_HIP(launch_copy(X + (size_t)idx[j] * n,
                                            T + (size_t)j * n, n, st));
                    B2G_HIP(launch_copy(T, X, (size_t)m * n, st));
                }
            }
            // one Gram-Schmidt pass of q against the subspace
            for (int j = 0; j < m; j++) {
                B2G_HIP(launch_dot(bv(j), w.q, n, ctx->red, w.scal, st));
                B2G_HIP(launch_axpy_dev(w.scal, -1.0, bv(j), w.q, n, st));
            }
            if ((r = project(w.q)))
                return r;
            double nq;
            if ((r = dev_dot(ctx, w.q, w.q, n, nq)))
                return r;
            B2G_HIP(launch_scale(1.0 / std::sqrt(nq), w.q, n, st));
            B2G_HIP(launch_copy(w.q, bv(m), n, st));
            m++;
        }
        if (xiter == o.soft_max_iter) {
            soft_stop = true;
            break;
        }
    }
    for (int i = 0; i < k; i++)
        B2G_HIP(launch_copy(bv(idx[i]), x + (size_t)i * n, n, st));
    B2G_HIP(hipStreamSynchronize(st));
    ndav = xiter;
    if (!converged && !soft_stop) {
        set_error("Davidson: only " + std::to_string(ck) + " converged");
        return B2G_ERR_NOCONV;
    }
    return B2G_OK;
}

// Graph-captured Davidson for k roots of a Normal-type solve: the first
// m - 1 products run eagerly as one multi-RHS launch, then every iteration
// replays one of two hipGraphs, the product of the newest vector plus
// launch_davidson_step, or the step alone after a step that appended no
// vector (a root converged), and the host reads a 128-byte status header
// from pinned memory after each replay.
static int davidson_graph(b2g_effham h, double *x, int k, const DavOpts &o,
                          int precision, std::vector<double> &eigvals,
                          int &ndav, uint64_t &nmv, bool &soft_stop) {
    b2g_ctx ctx = h->ctx;
    hipStream_t st = ctx->stream;
    const size_t n = h->prob.ket.total();
    B2G_REQUIRE(h->prob.bra.total() == n, B2G_ERR_ARG,
                "Davidson needs bra and ket of equal size");
    B2G_REQUIRE(k >= 1 && k <= 32, B2G_ERR_ARG, "bad number of roots");
    int def_min = std::max(o.def_min, k), def_max = o.def_max;
    if (def_max < k + k / 2)
        def_max = k + k / 2;
    if (def_max < 2)
        def_max = 2;
    B2G_REQUIRE(def_max <= 63 && def_min < def_max, B2G_ERR_ARG,
                "graph Davidson: need k <= def_min < def_max <= 63");
    B2G_REQUIRE(!h->plan->timing, B2G_ERR_STATE,
                "graph Davidson: disable plan timing first");
    // the subspace holds def_max vectors plus the one appended before the
    // restart check of the next iteration
    const size_t nv = (size_t)def_max + 1;
    SetupTimer tm("davidson");
    DavCache &c = h->dav;
    if (c.n != n || c.nv != nv) {
        B2G_HIP(hipStreamSynchronize(st));
        c.release();
        auto fail = [&c](hipError_t e) {
            if (e != hipSuccess)
                c.release();
            return e;
        };
        B2G_HIP(fail(hipMalloc(&c.B, nv * n * sizeof(double))));
        B2G_HIP(fail(hipMalloc(&c.S, nv * n * sizeof(double))));
        B2G_HIP(fail(hipMalloc(&c.q, n * sizeof(double))));
        B2G_HIP(fail(hipMalloc(&c.t, n * sizeof(double))));
        B2G_HIP(fail(hipMalloc(&c.qn, n * sizeof(double))));
        B2G_HIP(fail(hipMalloc(&c.sn, n * sizeof(double))));
        B2G_HIP(fail(hipMalloc(&c.partial,
                               davidson_partial_len() * sizeof(double))));
        B2G_HIP(fail(hipMalloc(&c.dst, sizeof(DavState))));
        B2G_HIP(fail(hipHostMalloc(&c.hdr, sizeof(DavState))));
        c.n = n;
        c.nv = nv;
        tm.mark("workspace");
    }
    // the initial vectors, Gram-Schmidt normalised (:927-955); dependent
    // ones cut the subspace to the independent leading ones
    auto bv = [&](int i) { return c.B + (size_t)i * n; };
    int r;
    for (int i = 0; i < k; i++)
        B2G_HIP(launch_copy(x + (size_t)i * n, bv(i), n, st));
    int m = k;
    for (int i = 0; i < k; i++) {
        for (int j = 0; j < i; j++) {
            double d;
            if ((r = dev_dot(ctx, bv(j), bv(i), n, d)))
                return r;
            B2G_HIP(launch_axpy(-d, bv(j), bv(i), n, st));
        }
        double nrm2;
        if ((r = dev_dot(ctx, bv(i), bv(i), n, nrm2)))
            return r;
        if (std::abs(nrm2) < 1E-14 && i > 0) {
            m = i;
            break;
        }
        B2G_REQUIRE(std::abs(nrm2) >= 1E-14, B2G_ERR_ARG,
                    "Davidson initial guess has zero norm");
        B2G_HIP(launch_scale(1.0 / std::sqrt(nrm2), bv(i), n, st));
    }
    if ((r = plan_reserve(h->plan, std::max(1, m - 1), precision)) != B2G_OK)
        return r;
    nmv = 0;
    if (m > 1) {
        if ((r = b2g_hpsi(ctx, h->plan, c.B, c.S, m - 1, 1.0, 0, precision)))
            return r;
        nmv += m - 1;
    }
    B2G_HIP(launch_copy(bv(m - 1), c.qn, n, st));
    DavState init;
    memset(&init, 0, sizeof init);
    init.m = m;
    init.max_iter = o.max_iter;
    init.soft_max_iter = o.soft_max_iter;
    init.def_min = def_min;
    init.def_max = def_max;
    init.k = k;
    init.appended = 1;
    init.type = o.type & (B2G_DAV_DAVIDSON_PRECOND | B2G_DAV_NO_PRECOND);
    init.conv = o.conv;
    init.rel2 = o.rel * o.rel;
    B2G_HIP(hipMemcpyAsync(c.dst, &init, sizeof init, hipMemcpyHostToDevice,
                           st));
    if ((r = plan_reserve(h->plan, 1, precision)) != B2G_OK)
        return r;
    B2G_HIP(hipStreamSynchronize(st));
    tm.mark("plan reserve");
    for (int with_product = 1; with_product >= 0; with_product--) {
        hipGraphExec_t &exec = c.exec[precision][with_product];
        if (exec != nullptr && c.gen[precision] == h->plan->scratch_gen)
            continue;
        if (exec) {
            hipGraphExecDestroy(exec);
            exec = nullptr;
        }
        // capture one iteration: (H qn -> sn and) the device step, which
        // ends by writing the state header to pinned host memory
        hipGraph_t graph = nullptr;
        B2G_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = with_product
                     ? b2g_hpsi(ctx, h->plan, c.qn, c.sn, 1, 1.0, 0, precision)
                     : B2G_OK;
        hipError_t e = rc == B2G_OK
                           ? launch_davidson_step(c.dst, c.B, c.S, c.sn, c.q,
                                                  c.t, c.qn, h->diag, n,
                                                  (int)nv, c.partial, c.hdr,
                                                  st)
                           : hipSuccess;
        const hipError_t e2 = hipStreamEndCapture(st, &graph);
        if (rc == B2G_OK && e == hipSuccess && e2 == hipSuccess)
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (graph)
            hipGraphDestroy(graph);
        if (rc != B2G_OK)
            return rc;
        B2G_HIP(e2);
        B2G_HIP(e);
        tm.mark(with_product ? "graph capture" : "graph capture (step)");
    }
    c.gen[precision] = h->plan->scratch_gen;
    DavState *hdr = c.hdr;
    ndav = 0;
    int status = 0;
    bool product = true;
    while (status == 0) {
        B2G_HIP(hipGraphLaunch(c.exec[precision][product ? 1 : 0], st));
        B2G_HIP(hipStreamSynchronize(st));
        if (product)
            nmv++;
        status = hdr->status;
        ndav = hdr->iter;
        product = hdr->appended != 0;
    }
    tm.mark("iterations");
    // energies and vectors of the k lowest Ritz pairs (roots past the
    // converged ones report 0 after a soft stop, as the reference)
    std::vector<double> ld(k);
    B2G_HIP(hipMemcpyAsync(ld.data(), c.dst->ld, k * sizeof(double),
                           hipMemcpyDeviceToHost, st));
    B2G_HIP(launch_copy(c.B, x, (size_t)k * n, st));
    B2G_HIP(hipStreamSynchronize(st));
    // roots with a value: up to ck, or below ck when the last step
    // converged a root (it appended no vector)
    const int nval = hdr->appended ? hdr->ck + 1 : hdr->ck;
    eigvals.assign(k, 0.0);
    for (int i = 0; i < k && i < nval; i++)
        eigvals[i] = ld[i];
    soft_stop = status == 3;
    if (status == 1 || status == 3)
        return B2G_OK;
    set_error("Davidson (graph): not converged after " +
              std::to_string(ndav) + " iterations");
    return B2G_ERR_NOCONV;
}

} // namespace b2g

namespace b2g {

// EffectiveHamiltonian::eigs with DavidsonTypes::Exact
// (IterativeMatrixFunctions::exact_diagonalization,
// src/core/iterative_matrix_functions.hpp:110-209): the dense matrix from
// products of unit vectors (multi-RHS launches), the device eigensolver,
// then the k targeted eigenpairs.  ndav = n, as the reference.
static int exact_diag(b2g_effham h, double *x, int k, const DavOpts &o,
                      int precision, std::vector<double> &eigvals, int &ndav,
                      uint64_t &nmv) {
    b2g_ctx ctx = h->ctx;
    hipStream_t st = ctx->stream;
    const size_t n = h->prob.ket.total();
    B2G_REQUIRE(h->prob.bra.total() == n, B2G_ERR_ARG,
                "exact diagonalisation needs bra and ket of equal size");
    B2G_REQUIRE(n >= 1 && n <= 16352 && (size_t)k <= n, B2G_ERR_ARG,
                "exact diagonalisation: order above 16352 or below k");
    B2G_REQUIRE(h->kernel == nullptr, B2G_ERR_STATE,
                "exact diagonalisation with a matvec hook is not supported");
    double *E = nullptr, *Hm = nullptr, *w = nullptr;
    B2G_HIP(hipMalloc(&Hm, n * n * sizeof(double)));
    std::unique_ptr<double, int (*)(void *)> g1(
        Hm, [](void *p) { return (int)hipFree(p); });
    B2G_HIP(hipMalloc(&E, n * n * sizeof(double)));
    std::unique_ptr<double, int (*)(void *)> g2(
        E, [](void *p) { return (int)hipFree(p); });
    B2G_HIP(hipMalloc(&w, n * sizeof(double)));
    std::unique_ptr<double, int (*)(void *)> g3(
        w, [](void *p) { return (int)hipFree(p); });
    // rows of the identity, 64 at a time (staged in E), their products as
    // rows of Hm (H is symmetric: row i of Hm = (H e_i)^T = column i)
    const size_t batch = 64;
    std::vector<double> eye(batch * n);
    for (size_t i = 0; i < n; i += batch) {
        const size_t nb = std::min(batch, n - i);
        std::fill(eye.begin(), eye.end(), 0.0);
        for (size_t j = 0; j < nb; j++)
            eye[j * n + i + j] = 1.0;
        B2G_HIP(hipMemcpyAsync(E, eye.data(), nb * n * sizeof(double),
                               hipMemcpyHostToDevice, st));
        int r = b2g_hpsi(ctx, h->plan, E, Hm + i * n, (int)nb, 1.0, 0,
                         precision);
        if (r)
            return r;
        B2G_HIP(hipStreamSynchronize(st));
    }
    nmv = n;
    B2G_HIP(hipStreamSynchronize(st));
    const uint32_t nn = (uint32_t)n;
    const uint64_t off = 0;
    int r = b2g_syev_batched(ctx, 1, &nn, &off, Hm, w, E, nullptr);
    if (r)
        return r;
    std::vector<double> ld(n);
    B2G_HIP(hipMemcpy(ld.data(), w, n * sizeof(double),
                      hipMemcpyDeviceToHost));
    std::vector<int> idx;
    ritz_order(o.type, o.shift, ld, (int)n, idx);
    eigvals.assign(k, 0.0);
    for (int i = 0; i < k; i++) {
        eigvals[i] = ld[idx[i]];
        B2G_HIP(launch_copy(E + (size_t)idx[i] * n, x + (size_t)i * n, n,
                            st));
    }
    B2G_HIP(hipStreamSynchronize(st));
    ndav = (int)n;
    return B2G_OK;
}

} // namespace b2g

extern "C" void b2g_davidson_opts_default(b2g_davidson_opts *o) {
    if (!o)
        return;
    memset(o, 0, sizeof *o);
    o->conv_thrd = 5E-6;
    o->max_iter = 5000;
    o->soft_max_iter = -1;
    o->def_min = 2;
    o->def_max = 50;
    o->davidson_type = B2G_DAV_NORMAL;
    o->precision = B2G_PREC_F64;
    o->refine = 1;
}

extern "C" int b2g_effham_eigs_ex(b2g_effham h, double *x, int k,
                                  const b2g_davidson_opts *opts,
                                  double *energies, int *ndav,
                                  uint64_t *nflop, double *tdav) {
    B2G_REQUIRE(h && x && energies && opts, B2G_ERR_ARG, "NULL arg");
    B2G_REQUIRE(h->plan, B2G_ERR_STATE, "effham was built host-only");
    const int precision = opts->precision;
    B2G_REQUIRE(precision == B2G_PREC_F64 || precision == B2G_PREC_F32 ||
                    precision == B2G_PREC_BF16X3,
                B2G_ERR_ARG, "unsupported precision");
    const int known = B2G_DAV_GREATER_THAN | B2G_DAV_LESS_THAN |
                      B2G_DAV_CLOSE_TO | B2G_DAV_DAVIDSON_PRECOND |
                      B2G_DAV_NO_PRECOND | B2G_DAV_EXACT |
                      B2G_DAV_ELEMENT_PROJ;
    B2G_REQUIRE((opts->davidson_type & ~known) == 0, B2G_ERR_ARG,
                "unsupported Davidson type (Harmonic, NonHermitian and "
                "LeftEigen are not built)");
    B2G_REQUIRE(opts->max_iter >= 1 && opts->soft_max_iter >= -1,
                B2G_ERR_ARG, "bad iteration limits");
    DavOpts o;
    o.conv = opts->conv_thrd;
    o.rel = opts->rel_conv_thrd;
    o.shift = opts->shift;
    o.max_iter = opts->max_iter;
    o.soft_max_iter = opts->soft_max_iter;
    o.def_min = opts->def_min;
    o.def_max = opts->def_max;
    o.type = opts->davidson_type;
    o.n_ortho = opts->n_ortho;
    o.ortho = opts->ortho;
    o.proj_weights = opts->proj_weights;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<double> ev;
    int nd = 0, nd2 = 0;
    uint64_t nmv = 0, nmv2 = 0;
    bool soft = false, soft2 = false;
    const bool exact = (o.type & B2G_DAV_EXACT) != 0;
    const bool graph = !exact && h->dav_mode == 1 && !h->plan->timing &&
                       h->kernel == nullptr && o.n_ortho == 0 &&
                       !(o.type & (B2G_DAV_GREATER_THAN | B2G_DAV_LESS_THAN |
                                   B2G_DAV_CLOSE_TO));
    auto solve = [&](const DavOpts &oo, int prec, int &ndv, uint64_t &nm,
                     bool &sf) {
        if (exact)
            return exact_diag(h, x, k, oo, prec, ev, ndv, nm);
        return graph ? davidson_graph(h, x, k, oo, prec, ev, ndv, nm, sf)
                     : davidson(h, x, k, oo, prec, ev, ndv, nm, sf);
    };
    // A reduced-precision phase followed by the FP64 refinement stops at the
    // residual its products resolve: with split-BF16 (FP32) products the
    // computed H x carries a relative error of a few 1e-6 (1e-7), so its
    // iterations past |r| ~ 1e-5 |lambda| (4e-7 |lambda|) converge to the
    // eigenvector of the rounded operator, which the refinement has to undo
    // anyway.  B2G_MIXED_REL overrides the relative threshold (0: the
    // caller's thresholds throughout).
    DavOpts o1 = o;
    if (precision != B2G_PREC_F64 && opts->refine) {
        double kappa = precision == B2G_PREC_BF16X3 ? 1e-5 : 4e-7;
        if (const char *e = getenv("B2G_MIXED_REL"))
            kappa = atof(e);
        o1.rel = std::max(o.rel, kappa);
    }
    int r = solve(o1, precision, nd, nmv, soft);
    if (r == B2G_OK && precision != B2G_PREC_F64 && opts->refine && !soft)
        r = solve(o, B2G_PREC_F64, nd2, nmv2, soft2);
    auto t1 = std::chrono::steady_clock::now();
    for (int i = 0; i < k && i < (int)ev.size(); i++)
        energies[i] = ev[i];
    if (ndav)
        *ndav = nd + nd2;
    if (nflop)
        *nflop = h->macs * (nmv + nmv2);
    if (tdav)
        *tdav = std::chrono::duration<double>(t1 - t0).count();
    return r;
}

extern "C" int b2g_effham_eigs(b2g_effham h, double *x, int k,
                               double conv_thrd, int max_iter, int def_min,
                               int def_max, int precision, int refine,
                               double *energies, int *ndav, uint64_t *nflop,
                               double *tdav) {
    b2g_davidson_opts o;
    b2g_davidson_opts_default(&o);
    o.conv_thrd = conv_thrd;
    o.max_iter = max_iter;
    o.def_min = def_min;
    o.def_max = def_max;
    o.precision = precision;
    o.refine = refine;
    return b2g_effham_eigs_ex(h, x, k, &o, energies, ndav, nflop, tdav);
}
