// Host-side plan builder (connection lists and pair recording).
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <sstream>

namespace b2g {

// magnitude below which a term or coupling factor is dropped, as TINY in the
// reference (src/core/csr_matrix.hpp:32)
static constexpr double kTiny = 1E-20;

int BlockInfo::find_state(uint64_t q) const {
    auto it = std::lower_bound(quanta.begin(), quanta.end(), q);
    if (it == quanta.end() || *it != q)
        return -1;
    return (int)(it - quanta.begin());
}

bool Connections::find_class(uint8_t conj, uint64_t abdq, int &begin,
                             int &end) const {
    if (conj > 3 || n[4] <= 0)
        return false;
    auto lo = quanta.begin() + n[conj], hi = quanta.begin() + n[conj + 1];
    auto it = std::lower_bound(lo, hi, abdq);
    if (it == hi || *it != abdq)
        return false;
    int ik = (int)(it - quanta.begin());
    begin = (int)idx[ik];
    end = ik == n[4] - 1 ? nc() : (int)idx[ik + 1];
    return true;
}

std::vector<SubLabel> uniq_sub_labels(const Problem &p) {
    const Symm s = p.symm;
    const uint64_t l = p.opdq;
    std::vector<SubLabel> r;
    for (const Term &t : p.terms) {
        uint64_t adq = p.lops[t.a].dq, bdq = p.rops[t.b].dq;
        uint64_t bra = (t.conj & 1) ? q_neg(s, adq) : adq;
        uint64_t ket = (t.conj & 2) ? bdq : q_neg(s, bdq);
        uint64_t q = q_combine(s, l, bra, ket);
        if (q != QINVALID)
            r.emplace_back(t.conj, q);
    }
    // the (vacuum x vacuum) class is always present
    uint64_t p0 = q_combine(s, l, p.hop_left_vacuum, q_at(s, q_sub(s, l, l), 0));
    if (p0 != QINVALID)
        r.emplace_back((uint8_t)0, p0);
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    return r;
}

static const BlockInfo *lookup(const std::vector<InfoRef> &infos,
                               uint64_t dq) {
    auto it = std::lower_bound(
        infos.begin(), infos.end(), dq,
        [](const InfoRef &a, uint64_t q) { return a.first < q; });
    if (it == infos.end() || it->first != dq)
        return nullptr;
    return it->second;
}

static void finish_classes(Connections &ci, const std::vector<SubLabel> &subdq) {
    ci.n[4] = (int)subdq.size();
    for (int i = 3; i >= 0; i--)
        if (ci.n[i] == -1)
            ci.n[i] = ci.n[i + 1];
    ci.quanta.resize(subdq.size());
    for (size_t i = 0; i < subdq.size(); i++)
        ci.quanta[i] = subdq[i].second;
}

void build_wfn_connections(Connections &ci, Symm s, const SU2Coupling &cg,
                           uint64_t cdq, uint64_t vdq, uint64_t opdq,
                           const std::vector<SubLabel> &subdq,
                           const std::vector<InfoRef> &ainfos,
                           const std::vector<InfoRef> &binfos,
                           const BlockInfo &cinfo, const BlockInfo &vinfo) {
    ci = Connections();
    if (ainfos.empty() || binfos.empty()) {
        ci.n[0] = ci.n[1] = ci.n[2] = ci.n[3] = ci.n[4] = 0;
        return;
    }
    const bool su2 = s == Symm::SU2;
    struct Ent {
        double f;
        uint64_t iv;
        uint32_t ia, ib, ic;
    };
    ci.idx.resize(subdq.size());
    for (size_t k = 0; k < subdq.size(); k++) {
        const uint8_t conj = subdq[k].first;
        if (ci.n[conj] == -1)
            ci.n[conj] = (int)k;
        const bool cja = conj & 1, cjb = (conj & 2) >> 1;
        ci.idx[k] = (uint32_t)ci.ia.size();
        const uint64_t q = subdq[k].second;
        const uint64_t adq0 = q_get_bra(s, q, opdq), bdq0 = q_get_ket(s, q);
        const uint64_t adq = cja ? q_neg(s, adq0) : adq0;
        const uint64_t bdq = cjb ? bdq0 : q_neg(s, bdq0);
        const BlockInfo *ainfo = lookup(ainfos, adq);
        const BlockInfo *binfo = lookup(binfos, bdq);
        if (ainfo == nullptr || binfo == nullptr)
            continue;
        // entries grouped by their rank within the output block: rank 0 of
        // every output block first, then rank 1, ... (the order in which
        // the reference flattens its per-block lists)
        std::vector<std::vector<Ent>> ranks;
        for (int iv = 0; iv < vinfo.n(); iv++) {
            size_t ip = 0;
            const uint64_t lq = q_get_bra(s, vinfo.quanta[iv], vdq);
            const uint64_t rq = q_neg(s, q_get_ket(s, vinfo.quanta[iv]));
            const uint64_t rqps = cjb ? q_add(s, rq, bdq) : q_sub(s, rq, bdq);
            for (int r = 0; r < q_count(s, rqps); r++) {
                const uint64_t rqp = q_at(s, rqps, r);
                const int ib = binfo->find_state(
                    cjb ? q_combine(s, bdq, rqp, rq) : q_combine(s, bdq, rq, rqp));
                if (ib == -1)
                    continue;
                const uint64_t lqps = q_sub(s, cdq, rqp);
                for (int l = 0; l < q_count(s, lqps); l++) {
                    const uint64_t lqp = q_at(s, lqps, l);
                    const int ia = ainfo->find_state(
                        cja ? q_combine(s, adq, lqp, lq)
                            : q_combine(s, adq, lq, lqp));
                    const int ic = cinfo.find_state(
                        q_combine(s, cdq, lqp, q_neg(s, rqp)));
                    if (ia == -1 || ic == -1)
                        continue;
                    double f;
                    if (su2) {
                        const int m = q_multiplicity(s, cdq) *
                                      q_multiplicity(s, opdq) *
                                      q_multiplicity(s, lq) *
                                      q_multiplicity(s, rq);
                        long double x =
                            sqrtl((long double)m) *
                            cg.ninej(q_twos(s, lqp), q_twos(s, rqp),
                                     q_twos(s, cdq), q_twos(s, adq),
                                     q_twos(s, bdq), q_twos(s, opdq),
                                     q_twos(s, lq), q_twos(s, rq),
                                     q_twos(s, vdq));
                        if (binfo->is_fermion && q_is_fermion(lqp))
                            x = -x;
                        if (cja)
                            x *= cg.transpose(q_twos(s, adq), q_twos(s, lq),
                                              q_twos(s, lqp));
                        if (cjb)
                            x *= cg.transpose(q_twos(s, bdq), q_twos(s, rq),
                                              q_twos(s, rqp));
                        f = (double)x;
                    } else
                        f = (binfo->is_fermion && q_is_fermion(lqp)) ? -1.0
                                                                     : 1.0;
                    if (std::abs(f) < kTiny)
                        continue;
                    if (ranks.size() <= ip)
                        ranks.emplace_back();
                    ranks[ip].push_back(
                        Ent{f, (uint64_t)iv, (uint32_t)ia, (uint32_t)ib,
                            (uint32_t)ic});
                    ip++;
                }
            }
        }
        for (auto &rk : ranks)
            for (auto &e : rk) {
                ci.factor.push_back(e.f);
                ci.stride.push_back(e.iv);
                ci.ia.push_back(e.ia);
                ci.ib.push_back(e.ib);
                ci.ic.push_back(e.ic);
            }
    }
    finish_classes(ci, subdq);
}

void build_diag_connections(Connections &ci, Symm s, const SU2Coupling &cg,
                            uint64_t cdq, uint64_t opdq,
                            const std::vector<SubLabel> &subdq,
                            const std::vector<InfoRef> &ainfos,
                            const std::vector<InfoRef> &binfos,
                            const BlockInfo &cinfo) {
    ci = Connections();
    if (ainfos.empty() || binfos.empty()) {
        ci.n[0] = ci.n[1] = ci.n[2] = ci.n[3] = ci.n[4] = 0;
        return;
    }
    const bool su2 = s == Symm::SU2;
    ci.idx.resize(subdq.size());
    for (size_t k = 0; k < subdq.size(); k++) {
        const uint8_t conj = subdq[k].first;
        if (ci.n[conj] == -1)
            ci.n[conj] = (int)k;
        const bool cja = conj & 1, cjb = (conj & 2) >> 1;
        ci.idx[k] = (uint32_t)ci.ia.size();
        const uint64_t q = subdq[k].second;
        const uint64_t adq0 = q_get_bra(s, q, opdq), bdq0 = q_get_ket(s, q);
        const uint64_t adq = cja ? q_neg(s, adq0) : adq0;
        const uint64_t bdq = cjb ? bdq0 : q_neg(s, bdq0);
        // only operator pairs that can be block-diagonal contribute
        if (q_at(s, q_add(s, adq, bdq), 0) != q_at(s, q_sub(s, adq, adq), 0))
            continue;
        const BlockInfo *ainfo = lookup(ainfos, adq);
        const BlockInfo *binfo = lookup(binfos, bdq);
        if (ainfo == nullptr || binfo == nullptr)
            continue;
        for (int ic = 0; ic < cinfo.n(); ic++) {
            const uint64_t aq = q_get_bra(s, cinfo.quanta[ic], cdq);
            const uint64_t bq = q_neg(s, q_get_ket(s, cinfo.quanta[ic]));
            const int ia = ainfo->find_state(aq), ib = binfo->find_state(bq);
            if (ia == -1 || ib == -1 || aq != q_get_bra(s, aq, adq) ||
                bq != q_get_bra(s, bq, bdq))
                continue;
            double f;
            if (su2) {
                const int m = q_multiplicity(s, cdq) * q_multiplicity(s, opdq) *
                              q_multiplicity(s, aq) * q_multiplicity(s, bq);
                long double x =
                    sqrtl((long double)m) *
                    cg.ninej(q_twos(s, aq), q_twos(s, bq), q_twos(s, cdq),
                             q_twos(s, adq), q_twos(s, bdq), q_twos(s, opdq),
                             q_twos(s, aq), q_twos(s, bq), q_twos(s, cdq));
                if (cja)
                    x *= cg.transpose(q_twos(s, adq), q_twos(s, aq),
                                      q_twos(s, aq));
                if (cjb)
                    x *= cg.transpose(q_twos(s, bdq), q_twos(s, bq),
                                      q_twos(s, bq));
                if (binfo->is_fermion && q_is_fermion(aq))
                    x = -x;
                f = (double)x;
            } else
                f = (binfo->is_fermion && q_is_fermion(aq)) ? -1.0 : 1.0;
            if (std::abs(f) < kTiny)
                continue;
            ci.ia.push_back((uint32_t)ia);
            ci.ib.push_back((uint32_t)ib);
            ci.ic.push_back((uint32_t)ic);
            ci.factor.push_back(f);
            ci.stride.push_back(0);
        }
    }
    finish_classes(ci, subdq);
}

static bool term_class(const Problem &p, const Term &t, const Connections &ci,
                       int &b, int &e) {
    const Symm s = p.symm;
    uint64_t adq = p.lops[t.a].dq, bdq = p.rops[t.b].dq;
    uint64_t abdq = q_combine(s, p.opdq, (t.conj & 1) ? q_neg(s, adq) : adq,
                              (t.conj & 2) ? bdq : q_neg(s, bdq));
    return ci.find_class(t.conj, abdq, b, e);
}

// The delayed operator's tensor-product entries of a three-tensor string:
// the class (dconj, dcdq.combine(op(dl).q, -op(dr).q)) of the delayed
// operator's info (three_tensor_product_multiply, operator_functions.hpp:
// 576-600).  Returns false with a message if absent.
static bool delayed_class(const Problem &p, const Term &t,
                          const Connections *&dinfo, int &b, int &e,
                          double &dscale, std::string &err) {
    const Symm s = p.symm;
    const bool dleft = p.delayed_side == 1;
    const Operator &dc = dleft ? p.lops[t.a] : p.rops[t.b];
    const Operator &da = p.dlops[t.da], &db = p.drops[t.db];
    if ((size_t)dc.info >= p.tp.size() || p.tp[dc.info].nc() == 0 ||
        p.tp[dc.info].n[4] < 0) {
        err = "delayed string without a tensor-product connection list";
        return false;
    }
    dinfo = &p.tp[dc.info];
    const uint64_t q = q_combine(s, dc.dq, (t.dconj & 1) ? q_neg(s, da.dq) : da.dq,
                                 (t.dconj & 2) ? db.dq : q_neg(s, db.dq));
    if (!dinfo->find_class(t.dconj, q, b, e)) {
        err = "delayed string: class absent from the tensor-product list";
        return false;
    }
    dscale = da.factor * db.factor;
    return true;
}

// block (dims, element offset) of sub-operator o in its info
struct SubBlock {
    uint32_t m, n;
    uint64_t off;
};
static SubBlock sub_block(const std::vector<BlockInfo> &infos,
                          const Operator &o, uint32_t i) {
    const BlockInfo &bi = infos[o.info];
    return SubBlock{bi.nbra[i], bi.nket[i], o.data_off + bi.shift[i]};
}

// pairs of one three-tensor string (GMatrixFunctions::three_rotate,
// src/core/matrix_functions.hpp:1046-1112, re-expressed as one rotate on
// windows of C and V)
static bool record_three(const Problem &p, const Term &t, size_t it,
                         const Connections &wfn, int b, int e, double scale,
                         std::vector<PairRec> &out, std::string &err) {
    const Connections *dinfo = nullptr;
    int db_ = 0, de_ = 0;
    double dscale = 1.0;
    if (!delayed_class(p, t, dinfo, db_, de_, dscale, err))
        return false;
    scale *= dscale;
    if (std::abs(scale) < kTiny)
        return true;
    const bool dleft = p.delayed_side == 1;
    const Operator &la = p.lops[t.a], &rb = p.rops[t.b];
    const Operator &oda = p.dlops[t.da], &odb = p.drops[t.db];
    const BlockInfo &ai = p.linfos[la.info], &bi = p.rinfos[rb.info];
    auto fail = [&](const char *what, int il) {
        std::ostringstream ss;
        ss << "term " << it << " connection " << il << ": " << what;
        err = ss.str();
        return false;
    };
    for (int il = b; il < e; il++) {
        const uint32_t ia = wfn.ia[il], ib = wfn.ib[il], ic = wfn.ic[il];
        const uint64_t iv = wfn.stride[il];
        const uint32_t idc = dleft ? ia : ib;
        const uint32_t *lo = std::lower_bound(dinfo->ic.data() + db_,
                                              dinfo->ic.data() + de_, idc);
        for (int idl = (int)(lo - dinfo->ic.data()); idl < de_ &&
                                                     dinfo->ic[idl] == idc;
             idl++) {
            const SubBlock da = sub_block(p.dlinfos, oda, dinfo->ia[idl]);
            const SubBlock db = sub_block(p.drinfos, odb, dinfo->ib[idl]);
            const uint64_t stride = dinfo->stride[idl];
            const bool sa = da.m == 1 && da.n == 1, sb = db.m == 1 && db.n == 1;
            if (!sa && !sb)
                return fail("neither delayed sub-operator block is 1 x 1", il);
            PairRec r;
            r.kind = 1;
            r.alpha = scale * wfn.factor[il] * dinfo->factor[idl];
            r.scal_off = (int64_t)(sa ? da.off : db.off);
            const uint32_t cm_full = p.ket.nbra[ic], cn_full = p.ket.nket[ic];
            const uint32_t vm_full = p.bra.nbra[iv], vn_full = p.bra.nket[iv];
            if (dleft) {
                // c = (da x db)(.T) * C * ket(.T): rows of C and V
                const bool cb = t.conj & 1;
                const bool ca = ((t.dconj & 1) != 0) ^ cb;
                const bool cbb = ((t.dconj & 2) != 0) ^ cb;
                const uint32_t am = (ca ? da.m : da.n) * (cbb ? db.m : db.n);
                const uint32_t cm = (ca ? da.n : da.m) * (cbb ? db.n : db.m);
                const uint32_t bra_n = ai.nket[ia];
                const uint64_t ast = cb ? stride / bra_n : stride % bra_n;
                const uint64_t cst = cb ? stride % bra_n : stride / bra_n;
                const SubBlock x = sa ? db : da;
                const bool tx = sa ? cbb : ca;
                if (ast + am > cm_full || cst + cm > vm_full)
                    return fail("delayed window outside its block", il);
                if ((tx ? x.m != am || x.n != cm : x.m != cm || x.n != am))
                    return fail("delayed sub-operator shape does not chain", il);
                r.a_off = x.off;
                r.b_off = rb.data_off + bi.shift[ib];
                r.c_off = p.ket.shift[ic] + ast * cn_full;
                r.v_off = p.bra.shift[iv] + cst * vn_full;
                r.c_m = am, r.c_n = cn_full, r.v_m = cm, r.v_n = vn_full;
                r.conj = (tx ? 1u : 0u) | (t.conj & 2);
                const uint32_t bm = bi.nbra[ib], bn = bi.nket[ib];
                if (!((t.conj & 2) ? (bm == r.c_n && bn == r.v_n)
                                   : (bm == r.v_n && bn == r.c_n)))
                    return fail("block shapes do not chain", il);
            } else {
                // c = bra(.T) * C * (da x db)(.T): columns of C and V
                const bool ck = !(t.conj & 2);
                const bool ca = ((t.dconj & 1) != 0) ^ ck;
                const bool cbb = ((t.dconj & 2) != 0) ^ ck;
                const uint32_t kn = (ca ? da.m : da.n) * (cbb ? db.m : db.n);
                const uint32_t km = (ca ? da.n : da.m) * (cbb ? db.n : db.m);
                const uint32_t ket_n = bi.nket[ib];
                const uint64_t ast = ck ? stride % ket_n : stride / ket_n;
                const uint64_t cst = ck ? stride / ket_n : stride % ket_n;
                const SubBlock y = sa ? db : da;
                const bool ty = sa ? cbb : ca;
                if (ast + km > cn_full || cst + kn > vn_full)
                    return fail("delayed window outside its block", il);
                // work = C_win (c_m x km) * op(y) (km x kn)
                if ((ty ? y.m != kn || y.n != km : y.m != km || y.n != kn))
                    return fail("delayed sub-operator shape does not chain", il);
                r.a_off = la.data_off + ai.shift[ia];
                r.b_off = y.off;
                r.c_off = p.ket.shift[ic] + ast;
                r.v_off = p.bra.shift[iv] + cst;
                r.c_m = cm_full, r.c_n = km, r.v_m = vm_full, r.v_n = kn;
                r.c_ld = cn_full, r.v_ld = vn_full;
                // op(y) untransposed is the "C * B" form (conj bit 1)
                r.conj = (t.conj & 1) | (ty ? 0u : 2u);
                const uint32_t am = ai.nbra[ia], an = ai.nket[ia];
                if (!((t.conj & 1) ? (am == r.c_m && an == r.v_m)
                                   : (am == r.v_m && an == r.c_m)))
                    return fail("block shapes do not chain", il);
            }
            out.push_back(r);
        }
    }
    return true;
}

std::vector<PairRec> record_pairs(const Problem &p, const Connections &wfn,
                                  std::string &err) {
    std::vector<PairRec> out;
    for (size_t it = 0; it < p.terms.size(); it++) {
        const Term &t = p.terms[it];
        const Operator &la = p.lops[t.a], &rb = p.rops[t.b];
        const double scale = t.factor * la.factor * rb.factor;
        if (std::abs(scale) < kTiny)
            continue;
        if (t.three) {
            int b = 0, e = 0;
            if (!term_class(p, t, wfn, b, e)) {
                std::ostringstream ss;
                ss << "term " << it << " has no connection class";
                err = ss.str();
                return {};
            }
            if (!record_three(p, t, it, wfn, b, e, scale, out, err))
                return {};
            continue;
        }
        int b = 0, e = 0;
        if (!term_class(p, t, wfn, b, e)) {
            std::ostringstream ss;
            ss << "term " << it << " has no connection class";
            err = ss.str();
            return {};
        }
        const BlockInfo &ai = p.linfos[la.info], &bi = p.rinfos[rb.info];
        for (int il = b; il < e; il++) {
            const uint32_t ia = wfn.ia[il], ib = wfn.ib[il], ic = wfn.ic[il];
            const uint64_t iv = wfn.stride[il];
            PairRec r;
            r.a_off = la.data_off + ai.shift[ia];
            r.b_off = rb.data_off + bi.shift[ib];
            r.c_off = p.ket.shift[ic];
            r.v_off = p.bra.shift[iv];
            r.c_m = p.ket.nbra[ic], r.c_n = p.ket.nket[ic];
            r.v_m = p.bra.nbra[iv], r.v_n = p.bra.nket[iv];
            r.conj = t.conj;
            r.alpha = scale * wfn.factor[il];
            // shape consistency of op(A) (v_m x c_m) and op(B) (v_n x c_n)
            const uint32_t am = ai.nbra[ia], an = ai.nket[ia];
            const uint32_t bm = bi.nbra[ib], bn = bi.nket[ib];
            const bool aok = (t.conj & 1) ? (am == r.c_m && an == r.v_m)
                                          : (am == r.v_m && an == r.c_m);
            const bool bok = (t.conj & 2) ? (bm == r.c_n && bn == r.v_n)
                                          : (bm == r.v_n && bn == r.c_n);
            if (!aok || !bok) {
                std::ostringstream ss;
                ss << "term " << it << " connection " << il
                   << ": block shapes do not chain";
                err = ss.str();
                return {};
            }
            out.push_back(r);
        }
    }
    return out;
}

// diagonal of one three-tensor string (three_tensor_product_diagonal,
// operator_functions.hpp:248-328; the diagonal sub-blocks only,
// batch_gemm.hpp:516-565)
static bool record_three_diag(const Problem &p, const Term &t, size_t it,
                              const Connections &dg, int b, int e,
                              double scale, std::vector<DiagRec> &out,
                              std::string &err) {
    const Connections *dinfo = nullptr;
    int db_ = 0, de_ = 0;
    double dscale = 1.0;
    if (!delayed_class(p, t, dinfo, db_, de_, dscale, err))
        return false;
    scale *= dscale;
    if (std::abs(scale) < kTiny)
        return true;
    const bool dleft = p.delayed_side == 1;
    const Operator &la = p.lops[t.a], &rb = p.rops[t.b];
    const Operator &oda = p.dlops[t.da], &odb = p.drops[t.db];
    const BlockInfo &ai = p.linfos[la.info], &bi = p.rinfos[rb.info];
    for (int il = b; il < e; il++) {
        const uint32_t ia = dg.ia[il], ib = dg.ib[il], ic = dg.ic[il];
        const uint32_t idc = dleft ? ia : ib;
        const uint32_t dm = dleft ? ai.nbra[ia] : bi.nbra[ib];
        const uint32_t *lo = std::lower_bound(dinfo->ic.data() + db_,
                                              dinfo->ic.data() + de_, idc);
        for (int idl = (int)(lo - dinfo->ic.data()); idl < de_ &&
                                                     dinfo->ic[idl] == idc;
             idl++) {
            const uint64_t stride = dinfo->stride[idl];
            const uint64_t dstrm = stride / dm, dstrn = stride % dm;
            if (dstrm != dstrn)
                continue; // off-diagonal sub-block
            const SubBlock da = sub_block(p.dlinfos, oda, dinfo->ia[idl]);
            const SubBlock db = sub_block(p.drinfos, odb, dinfo->ib[idl]);
            const bool sa = da.m == 1 && da.n == 1, sb = db.m == 1 && db.n == 1;
            const SubBlock x = sa ? db : da;
            if ((!sa && !sb) || x.m != x.n) {
                std::ostringstream ss;
                ss << "term " << it << " diagonal connection " << il
                   << ": delayed sub-operators neither 1 x 1 nor square";
                err = ss.str();
                return false;
            }
            DiagRec r;
            r.d_off = p.ket.shift[ic];
            r.alpha = scale * dg.factor[il] * dinfo->factor[idl];
            r.scal_off = (int64_t)(sa ? da.off : db.off);
            if (dleft) {
                r.a_off = x.off, r.a_n = x.n;
                r.b_off = rb.data_off + bi.shift[ib], r.b_n = bi.nbra[ib];
                r.r0 = (uint32_t)dstrm, r.c0 = 0;
            } else {
                r.a_off = la.data_off + ai.shift[ia], r.a_n = ai.nbra[ia];
                r.b_off = x.off, r.b_n = x.n;
                r.r0 = 0, r.c0 = (uint32_t)dstrm;
            }
            if (r.r0 + r.a_n > p.ket.nbra[ic] || r.c0 + r.b_n > p.ket.nket[ic]) {
                std::ostringstream ss;
                ss << "term " << it << " diagonal connection " << il
                   << ": delayed window outside its block";
                err = ss.str();
                return false;
            }
            out.push_back(r);
        }
    }
    return true;
}

std::vector<DiagRec> record_diag(const Problem &p, const Connections &dg,
                                 std::string &err) {
    std::vector<DiagRec> out;
    for (size_t it = 0; it < p.terms.size(); it++) {
        const Term &t = p.terms[it];
        const Operator &la = p.lops[t.a], &rb = p.rops[t.b];
        const double scale = t.factor * la.factor * rb.factor;
        if (std::abs(scale) < kTiny)
            continue;
        int b = 0, e = 0;
        // a string whose class has no diagonal connections adds nothing
        if (!term_class(p, t, dg, b, e))
            continue;
        if (t.three) {
            if (!record_three_diag(p, t, it, dg, b, e, scale, out, err))
                return {};
            continue;
        }
        const BlockInfo &ai = p.linfos[la.info], &bi = p.rinfos[rb.info];
        for (int il = b; il < e; il++) {
            const uint32_t ia = dg.ia[il], ib = dg.ib[il], ic = dg.ic[il];
            DiagRec r;
            r.a_off = la.data_off + ai.shift[ia];
            r.b_off = rb.data_off + bi.shift[ib];
            r.d_off = p.ket.shift[ic];
            r.a_n = ai.nbra[ia];
            r.b_n = bi.nbra[ib];
            r.alpha = scale * dg.factor[il];
            if (ai.nket[ia] != r.a_n || bi.nket[ib] != r.b_n ||
                p.ket.nbra[ic] != r.a_n || p.ket.nket[ic] != r.b_n) {
                std::ostringstream ss;
                ss << "term " << it << " diagonal connection " << il
                   << ": non-square or mismatched blocks";
                err = ss.str();
                return {};
            }
            out.push_back(r);
        }
    }
    return out;
}

uint64_t pair_macs(const PairRec &r) {
    // work = C * op(B)^T: c_m x v_n with inner c_n; V += op(A) * work:
    // v_m x v_n with inner c_m
    return (uint64_t)r.c_m * r.c_n * r.v_n + (uint64_t)r.v_m * r.c_m * r.v_n;
}

} // namespace b2g
