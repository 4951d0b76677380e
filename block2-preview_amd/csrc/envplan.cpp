// Host-side plans of environment blocking: fused bases and tensor-product
// connection lists.
#include "envplan.hpp"

#include <algorithm>
#include <atomic>
#include <functional>
#include <cmath>
#include <limits>
#include <map>
#include <sstream>

#include "common.hpp"

namespace b2g {

// magnitude below which a coupling factor is dropped (TINY,
// src/core/csr_matrix.hpp:32)
static constexpr double kTinyTp = 1E-20;

int StateSet::find(uint64_t x) const {
    auto it = std::lower_bound(q.begin(), q.end(), x);
    if (it == q.end() || *it != x)
        return -1;
    return (int)(it - q.begin());
}

uint64_t StateSet::total() const {
    uint64_t t = 0;
    for (uint32_t v : n)
        t += v;
    return t;
}

bool state_connections(Symm s, const StateSet &a, const StateSet &b,
                       const StateSet &c, StateConn &out, std::string &err) {
    // pairs grouped by coupled quantum, each group in (a, b) order
    std::map<uint64_t, std::vector<std::pair<uint32_t, uint32_t>>> mp;
    for (int i = 0; i < a.size(); i++)
        for (int j = 0; j < b.size(); j++) {
            const uint64_t qc = q_add(s, a.q[i], b.q[j]);
            const int cnt = q_count(s, qc);
            for (int k = 0; k < cnt; k++)
                mp[q_at(s, qc, k)].emplace_back((uint32_t)i, (uint32_t)j);
        }
    out.acc.assign(1, 0);
    out.ij.clear();
    for (int ic = 0; ic < c.size(); ic++) {
        auto it = mp.find(c.q[ic]);
        if (it == mp.end()) {
            err = "fused basis holds a quantum that no pair couples to";
            return false;
        }
        out.ij.insert(out.ij.end(), it->second.begin(), it->second.end());
        out.acc.push_back((uint32_t)out.ij.size());
    }
    return true;
}

StateSet state_tensor_product(Symm s, const StateSet &a, const StateSet &b,
                              const StateSet &cref) {
    const uint64_t cap = std::numeric_limits<uint32_t>::max();
    StateSet c;
    c.q = cref.q;
    std::vector<uint64_t> n(c.q.size(), 0);
    for (int i = 0; i < a.size(); i++)
        for (int j = 0; j < b.size(); j++) {
            const uint64_t qc = q_add(s, a.q[i], b.q[j]);
            const int cnt = q_count(s, qc);
            for (int k = 0; k < cnt; k++) {
                const int ic = c.find(q_at(s, qc, k));
                if (ic != -1)
                    n[ic] = std::min(n[ic] + (uint64_t)a.n[i] * b.n[j], cap);
            }
        }
    StateSet r;
    for (size_t i = 0; i < c.q.size(); i++)
        if (n[i] != 0) {
            r.q.push_back(c.q[i]);
            r.n.push_back((uint32_t)n[i]);
        }
    return r;
}

static const BlockInfo *find_info(const std::vector<InfoRef> &infos,
                                  uint64_t dq) {
    auto it = std::lower_bound(
        infos.begin(), infos.end(), dq,
        [](const InfoRef &x, uint64_t q) { return x.first < q; });
    if (it == infos.end() || it->first != dq)
        return nullptr;
    return it->second;
}

void build_tp_connections(Connections &ci, Symm s, const SU2Coupling &cg,
                          uint64_t cdq, const std::vector<SubLabel> &subdq,
                          const StateSet &bra, const StateSet &ket,
                          const StateSet &bra_a, const StateSet &bra_b,
                          const StateSet &ket_a, const StateSet &ket_b,
                          const StateConn &bra_c, const StateConn &ket_c,
                          const std::vector<InfoRef> &ainfos,
                          const std::vector<InfoRef> &binfos,
                          const BlockInfo &cinfo) {
    ci = Connections();
    if (ainfos.empty() || binfos.empty()) {
        ci.n[0] = ci.n[1] = ci.n[2] = ci.n[3] = ci.n[4] = 0;
        return;
    }
    const bool su2 = s == Symm::SU2;
    for (size_t k = 0; k < subdq.size(); k++) {
        const uint8_t conj = subdq[k].first;
        if (ci.n[conj] == -1)
            ci.n[conj] = (int)k;
        const bool cja = conj & 1, cjb = (conj & 2) >> 1;
        ci.idx.push_back((uint32_t)ci.ia.size());
        const uint64_t sq = subdq[k].second;
        const uint64_t adq = cja ? q_neg(s, q_get_bra(s, sq, cdq))
                                 : q_get_bra(s, sq, cdq);
        const uint64_t bdq = cjb ? q_get_ket(s, sq) : q_neg(s, q_get_ket(s, sq));
        const BlockInfo *ainfo = find_info(ainfos, adq);
        const BlockInfo *binfo = find_info(binfos, bdq);
        if (ainfo == nullptr || binfo == nullptr)
            continue;
        for (int ic = 0; ic < cinfo.n(); ic++) {
            const uint64_t cq = q_get_bra(s, cinfo.quanta[ic], cdq);
            const uint64_t cqprime = q_get_ket(s, cinfo.quanta[ic]);
            const int ib = bra.find(cq), ik = ket.find(cqprime);
            if (ib < 0 || ik < 0)
                continue;
            uint64_t bra_stride = 0;
            for (uint32_t kb = bra_c.acc[ib]; kb < bra_c.acc[ib + 1]; kb++) {
                const uint32_t jba = bra_c.ij[kb].first,
                               jbb = bra_c.ij[kb].second;
                uint64_t ket_stride = 0;
                for (uint32_t kk = ket_c.acc[ik]; kk < ket_c.acc[ik + 1];
                     kk++) {
                    const uint32_t jka = ket_c.ij[kk].first,
                                   jkb = ket_c.ij[kk].second;
                    const uint64_t qa =
                        cja ? q_combine(s, adq, ket_a.q[jka], bra_a.q[jba])
                            : q_combine(s, adq, bra_a.q[jba], ket_a.q[jka]);
                    const uint64_t qb =
                        cjb ? q_combine(s, bdq, ket_b.q[jkb], bra_b.q[jbb])
                            : q_combine(s, bdq, bra_b.q[jbb], ket_b.q[jkb]);
                    if (qa != QINVALID && qb != QINVALID) {
                        const int xa = ainfo->find_state(qa),
                                  xb = binfo->find_state(qb);
                        if (xa != -1 && xb != -1) {
                            const uint64_t aq = bra_a.q[jba],
                                           aqp = ket_a.q[jka];
                            const uint64_t bq = bra_b.q[jbb],
                                           bqp = ket_b.q[jkb];
                            // evaluated with the reference's precision
                            // steps: double sqrt, long double 9j and
                            // transpose factors, rounded to double after
                            // each multiplication
                            double factor = 1.0;
                            if (su2) {
                                const double sq = std::sqrt(
                                    (double)(q_multiplicity(s, cqprime) *
                                             q_multiplicity(s, cdq) *
                                             q_multiplicity(s, aq) *
                                             q_multiplicity(s, bq)));
                                factor = (double)((long double)sq *
                                                  cg.ninej(q_twos(s, aqp),
                                                           q_twos(s, bqp),
                                                           q_twos(s, cqprime),
                                                           q_twos(s, adq),
                                                           q_twos(s, bdq),
                                                           q_twos(s, cdq),
                                                           q_twos(s, aq),
                                                           q_twos(s, bq),
                                                           q_twos(s, cq)));
                            }
                            if (binfo->is_fermion && q_is_fermion(aqp))
                                factor = -factor;
                            if (su2 && cja)
                                factor = (double)((long double)factor *
                                                  cg.transpose(q_twos(s, adq),
                                                               q_twos(s, aq),
                                                               q_twos(s, aqp)));
                            if (su2 && cjb)
                                factor = (double)((long double)factor *
                                                  cg.transpose(q_twos(s, bdq),
                                                               q_twos(s, bq),
                                                               q_twos(s, bqp)));
                            if (std::abs(factor) >= kTinyTp) {
                                ci.ia.push_back((uint32_t)xa);
                                ci.ib.push_back((uint32_t)xb);
                                ci.ic.push_back((uint32_t)ic);
                                ci.stride.push_back(bra_stride *
                                                        cinfo.nket[ic] +
                                                    ket_stride);
                                ci.factor.push_back(factor);
                            }
                        }
                    }
                    ket_stride += (uint64_t)ket_a.n[jka] * ket_b.n[jkb];
                }
                bra_stride += (uint64_t)bra_a.n[jba] * bra_b.n[jbb];
            }
        }
    }
    ci.n[4] = (int)subdq.size();
    for (int i = 3; i >= 0; i--)
        if (ci.n[i] == -1)
            ci.n[i] = ci.n[i + 1];
    ci.quanta.resize(subdq.size());
    for (size_t i = 0; i < subdq.size(); i++)
        ci.quanta[i] = subdq[i].second;
}

bool WinPlan::finish(WinItems &items,
                     std::string &err) {
    // runs of consecutive records with one window, each run's group
    // (windows in first-appearance order), then the records of every group
    // in item order: a run's records land at its group's start plus the
    // records of the group's earlier runs, so runs are copied independently
    // (on the host workers) into the order of a stable sort by group
    SetupTimer tm("window plan finish");
    // run starts, found on the host workers over slices of the items
    const size_t N = items.size();
    const size_t nj = plan_jobs(N, 1u << 20);
    std::vector<std::vector<uint32_t>> slice_runs(nj);
    {
        std::vector<std::function<void()>> jobs;
        for (size_t j = 0; j < nj; j++)
            jobs.emplace_back([&, j]() {
                const size_t lo = N * j / nj, hi = N * (j + 1) / nj;
                for (size_t i = lo; i < hi; i++) {
                    const WinGroup &g = items[i].first;
                    if (i > 0) {
                        const WinGroup &p = items[i - 1].first;
                        if (p.c_off == g.c_off && p.rows == g.rows &&
                            p.cols == g.cols && p.ldc == g.ldc)
                            continue;
                    }
                    slice_runs[j].push_back((uint32_t)i);
                }
            });
        run_host_jobs(jobs);
    }
    tm.mark("run scan");
    std::vector<uint32_t> run_at, run_gid; // first item and group of a run
    {
        size_t nr = 0;
        for (const auto &v : slice_runs)
            nr += v.size();
        run_at.reserve(nr);
        for (auto &v : slice_runs) {
            run_at.insert(run_at.end(), v.begin(), v.end());
            std::vector<uint32_t>().swap(v);
        }
    }
    // Windows by their start element, checked on the host workers (job j
    // takes the starts whose hash falls in its share): two runs starting at
    // one element are either one window met again (then the runs are
    // grouped through a map, below) or windows of different shapes (an
    // error: windows must be disjoint).  When every run is a window of its
    // own, the groups are the runs in order.
    const size_t nrun0 = run_at.size();
    // the runs' windows, gathered into one compact array on the workers
    std::vector<WinGroup> rw(nrun0);
    {
        const size_t nc = plan_jobs(nrun0, 1u << 18);
        std::vector<std::function<void()>> jobs;
        for (size_t j = 0; j < nc; j++)
            jobs.emplace_back([&, j, nc]() {
                for (size_t q = nrun0 * j / nc; q < nrun0 * (j + 1) / nc; q++)
                    rw[q] = items[run_at[q]].first;
            });
        run_host_jobs(jobs);
    }
    tm.mark("run windows gathered");
    std::atomic<int> repeated{0}, clash{0};
    // the first run of each run's window
    std::vector<uint32_t> run_first(nrun0);
    {
        const size_t nh = plan_jobs(nrun0, 1u << 18);
        std::vector<std::function<void()>> jobs;
        for (size_t j = 0; j < nh; j++)
            jobs.emplace_back([&, j, nh]() {
                FlatMap<uint64_t, uint32_t, U64Hash> seen(nrun0 / nh + 16);
                for (size_t q = 0; q < nrun0; q++) {
                    const WinGroup &g = rw[q];
                    const uint64_t h = g.c_off * 0x9E3779B97F4A7C15ull;
                    if ((h >> 40) % nh != j)
                        continue;
                    auto r = seen.emplace(g.c_off, (uint32_t)q);
                    run_first[q] = *r.first;
                    if (r.second)
                        continue;
                    const WinGroup &o = rw[*r.first];
                    if (o.rows == g.rows && o.cols == g.cols && o.ldc == g.ldc)
                        repeated = 1;
                    else
                        clash = 1;
                }
            });
        run_host_jobs(jobs);
    }
    tm.mark("window check");
    if (clash) {
        err = "two windows of different shapes start at one element";
        return false;
    }
    std::vector<WinGroup> gs;
    std::vector<uint32_t> cnt;
    run_gid.resize(nrun0);
    if (!repeated) {
        gs.swap(rw);
        cnt.resize(nrun0);
        for (size_t q = 0; q < nrun0; q++) {
            run_gid[q] = (uint32_t)q;
            cnt[q] = (q + 1 < nrun0 ? run_at[q + 1] : (uint32_t)N) - run_at[q];
        }
    } else {
        // groups in first-appearance order: a run that starts its window
        // opens a group, a later run of it joins that group
        for (size_t q = 0; q < nrun0; q++) {
            const uint32_t f = run_first[q];
            if (f == q) {
                run_gid[q] = (uint32_t)gs.size();
                gs.push_back(rw[q]);
                cnt.push_back(0);
            } else {
                run_gid[q] = run_gid[f];
            }
            cnt[run_gid[q]] +=
                (q + 1 < nrun0 ? run_at[q + 1] : (uint32_t)N) - run_at[q];
        }
    }
    tm.mark("runs and windows");
    std::vector<uint32_t> pos(gs.size() + 1, 0);
    for (uint32_t g = 0; g < gs.size(); g++)
        pos[g + 1] = pos[g] + cnt[g];
    groups.assign(gs.begin(), gs.end());
    for (uint32_t g = 0; g < gs.size(); g++) {
        groups[g].rec_begin = pos[g];
        groups[g].rec_end = pos[g + 1];
    }
    const size_t nrun = run_at.size();
    std::vector<uint32_t> run_dst(nrun);
    for (size_t q = 0; q < nrun; q++) {
        const uint32_t end =
            q + 1 < nrun ? run_at[q + 1] : (uint32_t)items.size();
        run_dst[q] = pos[run_gid[q]];
        pos[run_gid[q]] += end - run_at[q];
    }
    recs.resize(items.size());
    tm.mark("offsets and allocation");
    // runs split into jobs of about equal record counts; per job the
    // algorithmic element counts and buffer extents, combined in job order
    struct Part {
        uint64_t src = 0, a = 0, b = 0;
    };
    std::vector<Part> part(nj);
    std::vector<std::function<void()>> jobs;
    size_t q0 = 0;
    for (size_t j = 0; j < nj; j++) {
        const uint64_t target = (uint64_t)items.size() * (j + 1) / nj;
        size_t q1 = q0;
        while (q1 < nrun && (j + 1 == nj || run_at[q1] < target))
            q1++;
        jobs.emplace_back([&, j, q0, q1]() {
            Part &pt = part[j];
            for (size_t q = q0; q < q1; q++) {
                const uint32_t end =
                    q + 1 < nrun ? run_at[q + 1] : (uint32_t)items.size();
                WinRec *d = recs.data() + run_dst[q];
                for (uint32_t i = run_at[q]; i < end; i++) {
                    const WinRec &r = items[i].second;
                    *d++ = r;
                    pt.src += (uint64_t)r.am * r.an + (uint64_t)r.bm * r.bn;
                    pt.a = std::max(pt.a, r.a_off + (r.am == 0 ? 0 :
                                    (uint64_t)(r.am - 1) * r.lda + r.an));
                    pt.b = std::max(pt.b, r.b_off + (r.bm == 0 ? 0 :
                                    (uint64_t)(r.bm - 1) * r.ldb + r.bn));
                }
            }
        });
        q0 = q1;
    }
    run_host_jobs(jobs);
    tm.mark("copy");
    src_elems = dst_elems = 0;
    for (const Part &pt : part) {
        src_elems += pt.src;
        a_need = std::max(a_need, pt.a);
        b_need = std::max(b_need, pt.b);
    }
    for (const WinGroup &w : groups) {
        dst_elems += (uint64_t)w.rows * w.cols;
        c_len = std::max(c_len, w.c_off + (w.rows == 0 ? 0 :
                         (uint64_t)(w.rows - 1) * w.ldc + w.cols));
    }
    return true;
}

} // namespace b2g
