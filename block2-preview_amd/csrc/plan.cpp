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

std::vector<PairRec> record_pairs(const Problem &p, const Connections &wfn,
                                  std::string &err) {
    std::vector<PairRec> out;
    for (size_t it = 0; it < p.terms.size(); it++) {
        const Term &t = p.terms[it];
        const Operator &la = p.lops[t.a], &rb = p.rops[t.b];
        const double scale = t.factor * la.factor * rb.factor;
        if (std::abs(scale) < kTiny)
            continue;
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
