// Binding of libb2g.so into block2: the code a block2 maintainer adds so
// that DMRG::two_dot_eigs_and_perturb (src/dmrg/sweep_algorithm.hpp:
// 1214-1262) solves each two-site effective Hamiltonian on the GPU.
//
// Header-only.  Include it after block2's own headers (block2_core.hpp,
// block2_dmrg.hpp), with include/b2g.h on the include path, and link the
// program against libb2g.so.  Only functions exported by include/b2g.h are
// called.
//
//   b2g_block2::enabled()        B2G_ENABLE=1 in the environment
//   b2g_block2::thread_ctx()     one b2g context (one HIP stream) per host
//                                thread, on device B2G_DEVICE (default 0)
//   b2g_block2::ProblemPack<S>   EffectiveHamiltonian<S, double> -> the
//                                b2g_problem of b2g_effham_create
//   b2g_block2::eigs(h, ...)     drop-in for EffectiveHamiltonian::eigs
//                                (src/dmrg/effective_hamiltonian.hpp:480-568)
//                                returning block2's (E, ndav, nflop, tdav)
//
// The branch in EffectiveHamiltonian<S, FL, MPS<S, FL>>::eigs is
//
//   if (b2g_block2::enabled() && metric == nullptr && ortho_bra.empty() &&
//       is_same<FL, double>::value)
//       return b2g_block2::eigs<S>(*this, conv_thrd, max_iter,
//                                  deflation_min_size, deflation_max_size);
#pragma once

#include "b2g.h"

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

namespace b2g_block2 {

inline bool enabled() {
    const char *e = std::getenv("B2G_ENABLE");
    return e != nullptr && std::atoi(e) != 0;
}

// a non-B2G_OK status becomes block2's usual runtime_error
inline void check(int rc) {
    if (rc != B2G_OK)
        throw std::runtime_error(std::string("b2g: ") + b2g_last_error());
}

// One context per host thread, destroyed with the thread.  block2's OpenMP
// regions never enter the solver concurrently on one thread.
inline b2g_ctx thread_ctx() {
    struct Holder {
        b2g_ctx c = nullptr;
        ~Holder() {
            if (c != nullptr)
                b2g_ctx_destroy(c);
        }
    };
    static thread_local Holder h;
    if (h.c == nullptr) {
        const char *d = std::getenv("B2G_DEVICE");
        check(b2g_ctx_create(d != nullptr ? std::atoi(d) : 0, 0, &h.c));
    }
    return h.c;
}

// The inputs of EffectiveHamiltonian's constructor
// (src/dmrg/effective_hamiltonian.hpp:126-224) as a b2g_problem: bra/ket
// infos, left/right operator infos (sorted by delta quantum, as block2
// keeps them), the operators the H_eff strings reference with their data
// concatenated per side, and the strings of the OpSum in op->mat->data[0]
// (OpProduct: kind 0; OpSumProd, a delayed three-tensor string of a
// DelayedOperatorTensor: kind 1 with the delayed tensor's sub-operators and
// the tensor-product ConnectionInfo of the delayed side's infos).  The
// pack owns every array the problem points to; it must outlive
// b2g_effham_create (which copies what it keeps).
template <typename S> struct ProblemPack {
    typedef block2::EffectiveHamiltonian<S, double> EH;
    typedef block2::SparseMatrixInfo<S> SMI;
    typedef block2::SparseMatrix<S, double> SM;
    typedef block2::OperatorTensor<S, double> OT;
    typedef block2::DelayedOperatorTensor<S, double> DOT;
    typedef block2::OpExpr<S> Expr;
    static_assert(sizeof(S) == sizeof(uint64_t),
                  "labels must be the packed 64-bit SZ / SU2 LongLong types");

    b2g_problem pb;

    explicit ProblemPack(const EH &h) {
        using namespace block2;
        std::memset(&pb, 0, sizeof(pb));
        pb.symmetry = std::is_same<S, SU2>::value ? B2G_SU2 : B2G_SZ;
        pb.opdq = (uint64_t)h.opdq.data;
        pb.hop_left_vacuum = (uint64_t)h.hop_left_vacuum.data;
        pb.ket = info(h.ket->info);
        pb.bra = info(h.bra->info);
        for (auto &x : h.left_op_infos) {
            ldq.push_back((uint64_t)x.first.data);
            linfos.push_back(info(x.second));
        }
        for (auto &x : h.right_op_infos) {
            rdq.push_back((uint64_t)x.first.data);
            rinfos.push_back(info(x.second));
        }
        const shared_ptr<DelayedOperatorTensor<S, double>> &op = h.op;
        shared_ptr<Expr> expr = op->mat->data[0];
        if (expr->get_type() == OpTypes::ExprRef)
            expr = dynamic_pointer_cast<OpExprRef<S>>(expr)->op;
        std::vector<shared_ptr<OpProduct<S, double>>> strings;
        if (expr->get_type() == OpTypes::Sum)
            strings = dynamic_pointer_cast<OpSum<S, double>>(expr)->strings;
        else if (expr->get_type() == OpTypes::Prod ||
                 expr->get_type() == OpTypes::SumProd)
            strings.push_back(dynamic_pointer_cast<OpProduct<S, double>>(expr));
        else if (expr->get_type() != OpTypes::Zero)
            throw std::runtime_error("b2g: unsupported H_eff expression");
        const bool dleft = op->lopt->get_type() == OperatorTensorTypes::Delayed;
        const bool dright = op->ropt->get_type() == OperatorTensorTypes::Delayed;
        shared_ptr<DOT> dopt =
            dleft ? dynamic_pointer_cast<DOT>(op->lopt)
                  : dright ? dynamic_pointer_cast<DOT>(op->ropt) : nullptr;
        for (auto &s : strings) {
            b2g_term t;
            std::memset(&t, 0, sizeof(t));
            const bool prod = s->get_type() == OpTypes::Prod;
            // an OpProduct string may name an operator of a delayed tensor
            // that has an expression: it is formed here, as block2's
            // Prod dispatch forms it before each multiply
            // (TensorFunctions::tensor_product, tensor_functions.hpp)
            t.a = lset.get(op->lopt, s->a, prod ? h.tf : nullptr);
            t.b = rset.get(op->ropt, s->b, prod ? h.tf : nullptr);
            t.conj = s->conj;
            t.factor = s->factor;
            t.da = t.db = -1;
            if (s->get_type() == OpTypes::SumProd) {
                auto sp = dynamic_pointer_cast<OpSumProd<S, double>>(s);
                if (dopt == nullptr || sp->ops.size() != 2)
                    throw std::runtime_error(
                        "b2g: OpSumProd string without a two-factor delayed "
                        "tensor");
                t.kind = 1;
                t.da = dlset.get(dopt->lopt, sp->ops[0], nullptr);
                t.db = drset.get(dopt->ropt, sp->ops[1], nullptr);
                t.dconj = (uint32_t)((int)sp->conjs[0] |
                                     ((int)sp->conjs[1] << 1));
            } else if (!prod)
                throw std::runtime_error("b2g: unsupported H_eff string");
            terms.push_back(t);
        }
        lset.emit(h.left_op_infos, lops, ldata);
        rset.emit(h.right_op_infos, rops, rdata);
        if (dopt != nullptr) {
            // sub-operator data follow the delayed side's own operators
            std::vector<double> &host = dleft ? ldata : rdata;
            dlset.emit_own(*this, dlinfos, dlops, host);
            drset.emit_own(*this, drinfos, drops, host);
            const auto &side = dleft ? h.left_op_infos : h.right_op_infos;
            for (auto &x : side) {
                b2g_tp_conn c;
                std::memset(&c, 0, sizeof(c));
                auto ci = x.second->cinfo;
                for (int k = 0; k < 5; k++)
                    c.n[k] = ci != nullptr ? ci->n[k] : -1;
                if (ci != nullptr && ci->nc > 0) {
                    c.nc = (int32_t)ci->nc;
                    c.quanta = (const uint64_t *)ci->quanta;
                    c.idx = ci->idx;
                    c.ia = ci->ia;
                    c.ib = ci->ib;
                    c.ic = ci->ic;
                    c.stride = ci->stride;
                    c.factor = ci->factor;
                }
                tps.push_back(c);
            }
            pb.delayed_side = dleft ? 1 : 2;
            pb.n_dl_infos = (int32_t)dlinfos.size();
            pb.n_dr_infos = (int32_t)drinfos.size();
            pb.n_dl_ops = (int32_t)dlops.size();
            pb.n_dr_ops = (int32_t)drops.size();
            pb.dl_infos = dlinfos.data();
            pb.dr_infos = drinfos.data();
            pb.dl_ops = dlops.data();
            pb.dr_ops = drops.data();
            pb.tp = tps.data();
        }
        pb.n_left_infos = (int32_t)linfos.size();
        pb.n_right_infos = (int32_t)rinfos.size();
        pb.n_left_ops = (int32_t)lops.size();
        pb.n_right_ops = (int32_t)rops.size();
        pb.n_terms = (int32_t)terms.size();
        pb.left_info_dq = ldq.data();
        pb.right_info_dq = rdq.data();
        pb.left_infos = linfos.data();
        pb.right_infos = rinfos.data();
        pb.left_ops = lops.data();
        pb.right_ops = rops.data();
        pb.terms = terms.data();
        pb.left_data = ldata.data();
        pb.right_data = rdata.data();
        pb.left_len = ldata.size();
        pb.right_len = rdata.size();
    }

  private:
    // the index of an operator among those the strings reference
    struct OpSet {
        std::map<Expr *, int> ids;
        std::vector<shared_ptr<SM>> mats;
        int get(const shared_ptr<OT> &opt, const shared_ptr<Expr> &x,
                const shared_ptr<block2::TensorFunctions<S, double>> &tf) {
            using namespace block2;
            auto it = ids.find(x.get());
            if (it != ids.end())
                return it->second;
            if (opt->ops.count(x) == 0)
                throw std::runtime_error("b2g: string names an operator "
                                         "missing from its tensor");
            const int id = (int)mats.size();
            ids[x.get()] = id;
            shared_ptr<SM> m = opt->ops.at(x);
            if (m->data == nullptr && tf != nullptr &&
                opt->get_type() == OperatorTensorTypes::Delayed) {
                auto dopt = dynamic_pointer_cast<DOT>(opt);
                if (dopt->exprs.count(x)) {
                    auto mm = make_shared<SM>(
                        make_shared<VectorAllocator<double>>());
                    mm->allocate(m->info);
                    const pair<size_t, double> zf = dopt->exprs.at(x);
                    tf->tensor_product(dopt->mat->data[zf.first] * zf.second,
                                       dopt->lopt->ops, dopt->ropt->ops, mm);
                    mm->factor = m->factor;
                    m = mm;
                }
            }
            mats.push_back(m);
            return id;
        }
        // operators over the side's infos; a delayed operator never formed
        // has structure only (flags bit 0)
        void emit(const std::vector<pair<S, shared_ptr<SMI>>> &infos,
                  std::vector<b2g_operator> &out,
                  std::vector<double> &data) const {
            for (auto &m : mats) {
                b2g_operator o;
                std::memset(&o, 0, sizeof(o));
                o.delta_quantum = (uint64_t)m->info->delta_quantum.data;
                o.factor = m->factor;
                o.info = -1;
                for (size_t k = 0; k < infos.size(); k++)
                    if (infos[k].first == m->info->delta_quantum)
                        o.info = (int32_t)k;
                if (o.info < 0)
                    throw std::runtime_error("b2g: operator without an info");
                o.flags = m->data == nullptr ? 1 : 0;
                o.data_off = data.size();
                if (m->data != nullptr)
                    data.insert(data.end(), m->data,
                                m->data + m->total_memory);
                out.push_back(o);
            }
        }
        // sub-operators of a delayed tensor, each with its own info
        void emit_own(ProblemPack &pk, std::vector<b2g_sparse_info> &infos,
                      std::vector<b2g_operator> &out,
                      std::vector<double> &data) const {
            for (auto &m : mats) {
                b2g_operator o;
                std::memset(&o, 0, sizeof(o));
                o.delta_quantum = (uint64_t)m->info->delta_quantum.data;
                o.factor = m->factor;
                o.info = (int32_t)infos.size();
                infos.push_back(pk.info(m->info));
                if (m->data == nullptr)
                    throw std::runtime_error("b2g: delayed sub-operator "
                                             "without data");
                o.data_off = data.size();
                data.insert(data.end(), m->data, m->data + m->total_memory);
                out.push_back(o);
            }
        }
    };

    // SparseMatrixInfo -> b2g_sparse_info: labels pass through unchanged;
    // bond dimensions (ubond_t) and block shifts (uint32) are widened
    b2g_sparse_info info(const shared_ptr<SMI> &i) {
        keep32.emplace_back(i->n_states_bra, i->n_states_bra + i->n);
        const uint32_t *nb = keep32.back().data();
        keep32.emplace_back(i->n_states_ket, i->n_states_ket + i->n);
        const uint32_t *nk = keep32.back().data();
        keep64.emplace_back(i->block_shifts, i->block_shifts + i->n);
        b2g_sparse_info r;
        std::memset(&r, 0, sizeof(r));
        r.delta_quantum = (uint64_t)i->delta_quantum.data;
        r.is_fermion = i->is_fermion;
        r.n = i->n;
        r.quanta = (const uint64_t *)i->quanta;
        r.n_states_bra = nb;
        r.n_states_ket = nk;
        r.block_shifts = keep64.back().data();
        return r;
    }

    std::deque<std::vector<uint32_t>> keep32;
    std::deque<std::vector<uint64_t>> keep64;
    std::vector<b2g_sparse_info> linfos, rinfos, dlinfos, drinfos;
    std::vector<uint64_t> ldq, rdq;
    std::vector<b2g_operator> lops, rops, dlops, drops;
    std::vector<b2g_term> terms;
    std::vector<double> ldata, rdata;
    std::vector<b2g_tp_conn> tps;
    OpSet lset, rset, dlset, drset;
};

// EffectiveHamiltonian<S, double>::eigs on the GPU (one root, block2's ket
// as the initial vector and the result; src/dmrg/effective_hamiltonian.hpp:
// 480-568): (E, ndav, nflop, tdav) with nflop in block2's multiply-adds.
// precision B2G_PREC_F64, or B2G_PREC_F32 / B2G_PREC_BF16X3 with a final
// FP64 refinement.
template <typename S>
std::tuple<double, int, size_t, double>
eigs(block2::EffectiveHamiltonian<S, double> &h, double conv_thrd,
     int max_iter, int def_min, int def_max,
     int precision = B2G_PREC_F64) {
    ProblemPack<S> pk(h);
    b2g_ctx ctx = thread_ctx();
    b2g_effham eh = nullptr;
    check(b2g_effham_create(ctx, &pk.pb, precision == B2G_PREC_F32 ? 1u : 0u,
                            &eh));
    const size_t n = h.ket->total_memory;
    b2g_buf x = nullptr;
    double e = 0.0, tdav = 0.0;
    int ndav = 0;
    uint64_t nflop = 0;
    int rc = b2g_buf_create(ctx, n, B2G_F64, &x);
    if (rc == B2G_OK)
        rc = b2g_buf_write(x, h.ket->data, 0, n);
    if (rc == B2G_OK)
        rc = b2g_effham_eigs(eh, (double *)b2g_buf_device_ptr(x), 1,
                             conv_thrd, max_iter, def_min, def_max, precision,
                             precision != B2G_PREC_F64, &e, &ndav, &nflop,
                             &tdav);
    if (rc == B2G_OK)
        rc = b2g_buf_read(x, h.ket->data, 0, n);
    std::string msg = rc == B2G_OK ? "" : b2g_last_error();
    b2g_buf_destroy(x);
    b2g_effham_destroy(eh);
    if (rc != B2G_OK)
        throw std::runtime_error("b2g: " + msg);
    return std::make_tuple(e, ndav, (size_t)nflop, tdav);
}

} // namespace b2g_block2
