// Reference harness: TEST INFRASTRUCTURE ONLY.
//
// Compiled from the reference's own header-only sources (src/block2_core.hpp,
// src/block2_dmrg.hpp) by oracle/Makefile into oracle/_ref/b2ref.  It runs a
// plain two-site DMRG with the reference (MPOQC conventional MPO, as in
// unit_test/test_dmrg_n2_sto3g.cpp:53-70) and, at chosen (sweep, site) points,
// dumps the effective-Hamiltonian problem handed to
// EffectiveHamiltonian::eigs (src/dmrg/sweep_algorithm.hpp:1214-1250):
//   * ket/bra SparseMatrixInfo (src/core/sparse_matrix.hpp:47-877),
//   * left/right operator infos and operator data (lopt/ropt ops),
//   * the H_eff expression (OpSum of OpProduct, src/core/expr.hpp:503,811),
//   * the reference's ConnectionInfo lists for H*psi and for the diagonal
//     (sparse_matrix.hpp:81-292),
//   * psi, H*psi computed by the reference in SeqTypes::None mode
//     (operator_functions.hpp:474-542 -> matrix_functions.hpp:998-1009),
//   * the diagonal (effective_hamiltonian.hpp:190-202),
//   * a Davidson solve on that site (effective_hamiltonian.hpp:480-568),
//     its vector also in block2's sparse-matrix file format, and the
//     reference's density-matrix truncation of it (dump_split).
// Output goes to a directory of .npy files plus meta.txt; tests/golden/
// make_fixtures.py packs them into the committed .npz fixtures.
// Nothing in the product links or loads this program.  Built with
// -DB2G_SHIM_CHECK (oracle/_ref/b2shim) it also runs the block2 binding of
// libb2g.so on each dumped site (shim_check).

#include "block2_core.hpp"
#include "block2_dmrg.hpp"
#ifdef B2G_SHIM_CHECK
// the block2 binding of libb2g.so (built as oracle/_ref/b2shim)
#include "b2g_block2.hpp"
#endif

#include <chrono>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <sys/stat.h>

using namespace block2;

// ---------------------------------------------------------------- npy writer
template <typename T> static const char *npy_descr();
template <> const char *npy_descr<double>() { return "<f8"; }
template <> const char *npy_descr<uint64_t>() { return "<u8"; }
template <> const char *npy_descr<uint32_t>() { return "<u4"; }
template <> const char *npy_descr<int32_t>() { return "<i4"; }
template <> const char *npy_descr<uint8_t>() { return "|u1"; }

template <typename T>
static void write_npy(const string &path, const T *data, size_t n) {
    stringstream hs;
    hs << "{'descr': '" << npy_descr<T>() << "', 'fortran_order': False, "
       << "'shape': (" << n << ",), }";
    string h = hs.str();
    size_t total = 10 + h.size() + 1;
    size_t pad = (64 - total % 64) % 64;
    h += string(pad, ' ');
    h += '\n';
    ofstream ofs(path, ios::binary);
    const char magic[] = "\x93NUMPY";
    ofs.write(magic, 6);
    char ver[2] = {1, 0};
    ofs.write(ver, 2);
    uint16_t hl = (uint16_t)h.size();
    ofs.write((const char *)&hl, 2);
    ofs.write(h.data(), h.size());
    ofs.write((const char *)data, sizeof(T) * n);
}
template <typename T>
static void write_npy(const string &path, const vector<T> &v) {
    write_npy(path, v.data(), v.size());
}

template <typename T> static vector<T> read_npy(const string &path) {
    ifstream f(path, ios::binary);
    if (!f.good())
        throw runtime_error("cannot open " + path);
    char magic[6];
    uint8_t ver[2];
    f.read(magic, 6);
    f.read((char *)ver, 2);
    uint32_t hl = 0;
    if (ver[0] == 1) {
        uint16_t h = 0;
        f.read((char *)&h, 2);
        hl = h;
    } else
        f.read((char *)&hl, 4);
    string hdr(hl, ' ');
    f.read(&hdr[0], hl);
    if (hdr.find(npy_descr<T>()) == string::npos ||
        hdr.find("'fortran_order': False") == string::npos)
        throw runtime_error("unexpected dtype or order in " + path);
    const size_t start = (size_t)f.tellg();
    f.seekg(0, ios::end);
    const size_t n = ((size_t)f.tellg() - start) / sizeof(T);
    vector<T> v(n);
    f.seekg(start);
    f.read((char *)v.data(), n * sizeof(T));
    return v;
}

// --------------------------------------------------------------- settings
struct Settings {
    string fcidump, sym = "su2", outdir = "ref_out";
    vector<int> bdims{50};
    vector<double> noises{1E-5, 0.0};
    int nsweeps = 6, threads = 4, dump_sweep = -1;
    vector<int> dump_sites;
    double dav_conv = 1E-8, dump_conv = 1E-10;
    bool delayed = false;
    string noise_type = "dm";
    string load_dir;   // --load-problem: block2's H*psi on a given problem
    int reps = 3;
    // --save-sweep 1: block2's MPO and initial MPS written in its own file
    // formats (MPO::save_data, MPSInfo::save_data, SparseMatrix::save_data)
    // under OUT/sweep, with a record of every sweep step; --seed fixes the
    // generator for the random initial MPS (0: block2's clock seed) and
    // --noise-seed re-seeds it right before the sweeps (0: no re-seed)
    bool save_sweep = false;
    unsigned seed = 0, noise_seed = 0;
    // --dav-thrd: one Davidson threshold for every sweep (0: block2's
    // default, derived from the noises)
    double dav_thrd = 0.0;
};
static Settings G;

template <typename S> static uint64_t qdata(S q) { return (uint64_t)q.data; }

// ------------------------------------------------------------ info packing
template <typename S> struct InfoPack {
    vector<uint64_t> dq, quanta;
    vector<uint32_t> nbra, nket, shift, n, off, fermion;
    void add(const shared_ptr<SparseMatrixInfo<S>> &info) {
        dq.push_back(qdata(info->delta_quantum));
        fermion.push_back(info->is_fermion);
        n.push_back((uint32_t)info->n);
        off.push_back((uint32_t)quanta.size());
        for (int i = 0; i < info->n; i++) {
            quanta.push_back(qdata(info->quanta[i]));
            nbra.push_back((uint32_t)info->n_states_bra[i]);
            nket.push_back((uint32_t)info->n_states_ket[i]);
            shift.push_back(info->block_shifts[i]);
        }
    }
    void write(const string &d, const string &p) const {
        write_npy(d + "/" + p + "_dq.npy", dq);
        write_npy(d + "/" + p + "_fermion.npy", fermion);
        write_npy(d + "/" + p + "_n.npy", n);
        write_npy(d + "/" + p + "_off.npy", off);
        write_npy(d + "/" + p + "_quanta.npy", quanta);
        write_npy(d + "/" + p + "_nbra.npy", nbra);
        write_npy(d + "/" + p + "_nket.npy", nket);
        write_npy(d + "/" + p + "_shift.npy", shift);
    }
};

template <typename S, typename CI>
static void write_cinfo(const string &d, const string &p, const CI &ci) {
    vector<int32_t> n(ci->n, ci->n + 5);
    write_npy(d + "/" + p + "_n.npy", n);
    int nq = ci->n[4], nc = ci->nc;
    vector<uint64_t> q(nq);
    for (int i = 0; i < nq; i++)
        q[i] = qdata(ci->quanta[i]);
    write_npy(d + "/" + p + "_quanta.npy", q);
    write_npy(d + "/" + p + "_idx.npy", ci->idx, (size_t)nq);
    write_npy(d + "/" + p + "_stride.npy", ci->stride, (size_t)nc);
    write_npy(d + "/" + p + "_factor.npy", ci->factor, (size_t)nc);
    write_npy(d + "/" + p + "_ia.npy", ci->ia, (size_t)nc);
    write_npy(d + "/" + p + "_ib.npy", ci->ib, (size_t)nc);
    write_npy(d + "/" + p + "_ic.npy", ci->ic, (size_t)nc);
}

// operator set: index per referenced OpExpr, data concatenated
template <typename S> struct OpSet {
    map<OpExpr<S> *, int> ids;
    vector<shared_ptr<SparseMatrix<S, double>>> mats;
    // tf: when given, an operator of a delayed tensor that has an
    // expression (DelayedOperatorTensor::exprs) is formed from it the way
    // the reference's OpProduct dispatch forms it before each multiply
    // (tensor_functions.hpp, Prod case: tensor_product of the delayed
    // tensor's sub-operators)
    int get(const shared_ptr<OperatorTensor<S, double>> &opt,
            const shared_ptr<OpExpr<S>> &x,
            const shared_ptr<TensorFunctions<S, double>> &tf = nullptr) {
        auto it = ids.find(x.get());
        if (it != ids.end())
            return it->second;
        assert(opt->ops.count(x));
        int id = (int)mats.size();
        ids[x.get()] = id;
        shared_ptr<SparseMatrix<S, double>> m = opt->ops.at(x);
        if (m->data == nullptr && tf != nullptr &&
            opt->get_type() == OperatorTensorTypes::Delayed) {
            auto dopt = dynamic_pointer_cast<DelayedOperatorTensor<S, double>>(opt);
            if (dopt->exprs.count(x)) {
                auto mm = make_shared<SparseMatrix<S, double>>(
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
    void write(const string &d, const string &p,
               const vector<pair<S, shared_ptr<SparseMatrixInfo<S>>>> &infos)
        const {
        vector<uint64_t> dq, off;
        vector<double> factor, data;
        vector<int32_t> info_idx, flags;
        for (auto &m : mats) {
            dq.push_back(qdata(m->info->delta_quantum));
            factor.push_back(m->factor);
            int ii = -1;
            for (size_t k = 0; k < infos.size(); k++)
                if (infos[k].first == m->info->delta_quantum)
                    ii = (int)k;
            assert(ii != -1 && infos[ii].second->n == m->info->n);
            info_idx.push_back(ii);
            off.push_back(data.size());
            // a delayed operator has structure but no data
            flags.push_back(m->data == nullptr ? 1 : 0);
            if (m->data != nullptr)
                data.insert(data.end(), m->data, m->data + m->total_memory);
        }
        write_npy(d + "/" + p + "op_dq.npy", dq);
        write_npy(d + "/" + p + "op_factor.npy", factor);
        write_npy(d + "/" + p + "op_info.npy", info_idx);
        write_npy(d + "/" + p + "op_off.npy", off);
        write_npy(d + "/" + p + "op_flags.npy", flags);
        write_npy(d + "/" + p + "op_data.npy", data);
    }
    // sub-operators of a delayed tensor: each with its own info
    void write_own(const string &d, const string &p) const {
        InfoPack<S> ip;
        vector<uint64_t> dq, off;
        vector<double> factor, data;
        vector<int32_t> info_idx;
        for (auto &m : mats) {
            dq.push_back(qdata(m->info->delta_quantum));
            factor.push_back(m->factor);
            info_idx.push_back((int32_t)info_idx.size());
            ip.add(m->info);
            off.push_back(data.size());
            assert(m->data != nullptr);
            data.insert(data.end(), m->data, m->data + m->total_memory);
        }
        ip.write(d, p + "info");
        write_npy(d + "/" + p + "op_dq.npy", dq);
        write_npy(d + "/" + p + "op_factor.npy", factor);
        write_npy(d + "/" + p + "op_info.npy", info_idx);
        write_npy(d + "/" + p + "op_off.npy", off);
        write_npy(d + "/" + p + "op_data.npy", data);
    }
};

#ifdef B2G_SHIM_CHECK
// The block2 binding (block2-preview_amd/integration/b2g_block2.hpp) run on
// the live effective Hamiltonian: its b2g_problem, built into a host-only
// b2g effective Hamiltonian, must give the reference's own ConnectionInfo
// (wfn) entry by entry and the reference's multiply-add count.  Writes
// shim.txt: strings, OpSumProd strings, connections, entries equal, factor
// max relative difference, b2g MACs, reference MACs.
template <typename S>
static void shim_check(const shared_ptr<EffectiveHamiltonian<S, double>> &h,
                       const string &d, size_t ref_nflop) {
    b2g_block2::ProblemPack<S> pk(*h);
    b2g_effham eh = nullptr;
    b2g_block2::check(
        b2g_effham_create(nullptr, &pk.pb, B2G_EFFHAM_HOST_ONLY, &eh));
    uint64_t nc = 0;
    b2g_block2::check(b2g_effham_connections(eh, 0, &nc, nullptr, nullptr,
                                             nullptr, nullptr, nullptr));
    vector<uint32_t> ia(nc), ib(nc), ic(nc);
    vector<uint64_t> st(nc);
    vector<double> fa(nc);
    b2g_block2::check(b2g_effham_connections(eh, 0, &nc, ia.data(), ib.data(),
                                             ic.data(), st.data(), fa.data()));
    uint64_t kl = 0, bl = 0, np = 0, macs = 0;
    b2g_block2::check(b2g_effham_info(eh, &kl, &bl, &np, &macs));
    b2g_effham_destroy(eh);
    auto ci = h->wfn_infos[0];
    size_t equal = 0;
    double fdiff = 0.0;
    if ((size_t)ci->nc == nc)
        for (size_t k = 0; k < nc; k++) {
            equal += ia[k] == ci->ia[k] && ib[k] == ci->ib[k] &&
                     ic[k] == ci->ic[k] && st[k] == ci->stride[k];
            fdiff = max(fdiff, abs(fa[k] - ci->factor[k]) /
                                   max(abs(ci->factor[k]), 1e-300));
        }
    int nsp = 0;
    for (int t = 0; t < pk.pb.n_terms; t++)
        nsp += pk.pb.terms[t].kind == 1;
    ofstream f(d + "/shim.txt");
    f << setprecision(17) << pk.pb.n_terms << " " << nsp << " " << nc << " "
      << equal << " " << fdiff << " " << macs << " " << ref_nflop << "\n";
}
#endif

template <typename S> struct Dumper : CallbackKernel {
    DMRG<S, double, double> *dmrg = nullptr;
    mutable int ndumped = 0;
    // one record per sweep step (DMRG::sweep::iter.end): sweep, step,
    // the step's Davidson energy, discarded weight, the generator's next
    // output, and the bond basis the step left (total, checksum)
    mutable vector<double> steps;
    // wall-clock seconds of each sweep (DMRG::sweep.start .. sweep.end)
    mutable vector<double> sweep_s;
    mutable chrono::steady_clock::time_point sweep_t0;
    void compute(const string &name, int iprint) const override {
        if (name == "DMRG::sweep.start") {
            sweep_t0 = chrono::steady_clock::now();
            return;
        }
        if (name == "DMRG::sweep.end") {
            sweep_s.push_back(chrono::duration<double>(
                                  chrono::steady_clock::now() - sweep_t0)
                                  .count());
            return;
        }
        if (name == "DMRG::sweep::iter.end" && dmrg != nullptr &&
            dmrg->sweep_energies.size() != 0) {
            steps.push_back((double)dmrg->isweep);
            steps.push_back((double)(dmrg->sweep_energies.size() - 1));
            steps.push_back((double)dmrg->sweep_energies.back().back());
            steps.push_back((double)dmrg->sweep_discarded_weights.back());
            // the generator's next output, read from a copy (the stream
            // itself is left untouched): locates block2's position in its
            // noise stream after the step
            mt19937 peek = Random::rng();
            steps.push_back((double)peek());
            // the bond basis the step left behind: total states and a
            // checksum over (quantum, count)
            // the step's site: sweeps alternate starting forward from the
            // first site
            auto mi = dmrg->me->ket->info;
            const int st = (int)dmrg->sweep_energies.size() - 1;
            const bool fw = (dmrg->isweep % 2) == 0;
            const int ic = fw ? st : dmrg->me->n_sites - 2 - st;
            shared_ptr<StateInfo<S>> bs;
            if (fw) {
                mi->load_left_dims(ic + 1);
                bs = mi->left_dims[ic + 1];
            } else {
                mi->load_right_dims(ic + 1);
                bs = mi->right_dims[ic + 1];
            }
            double cks = 0;
            for (int q = 0; q < bs->n; q++)
                cks += (double)(bs->quanta[q].data % 1000003ULL) *
                       bs->n_states[q];
            steps.push_back((double)bs->n_states_total);
            steps.push_back(cks);
            return;
        }
        if (name != "DMRG::sweep::iter.eff_ham" || dmrg == nullptr)
            return;
        auto h_eff = dmrg->current_eff_ham;
        int isite = dmrg->me->center;
        if (G.dump_sweep != dmrg->isweep ||
            find(G.dump_sites.begin(), G.dump_sites.end(), isite) ==
                G.dump_sites.end())
            return;
        stringstream ss;
        ss << G.outdir << "/site" << isite << "_sw" << dmrg->isweep;
        string d = ss.str();
        mkdir(d.c_str(), 0755);
        dump(h_eff, d, isite);
        if (G.save_sweep)
            dump_environment(h_eff, d, isite);
        ndumped++;
    }
    // The environment step around a dumped site, in block2's own formats
    // (OperatorTensor::save_data, src/core/operator_tensor.hpp:122-148;
    // SparseMatrix::save_data with its info): the left / right environments
    // the effective Hamiltonian was blocked from (envs[center]->left /
    // right, src/dmrg/moving_environment.hpp:1745-2068), the blocked
    // operators of the effective Hamiltonian (lopt, ropt) and the MPS
    // tensors left of the center and right of the two sites (the rotations
    // that formed the environments, left_contract_rotate :226-,
    // right_contract_rotate :460-).
    void dump_environment(const shared_ptr<EffectiveHamiltonian<S, double>> &h,
                          const string &d, int isite) const {
        auto me = dmrg->me;
        auto save_t = [](const shared_ptr<OperatorTensor<S, double>> &t,
                         const string &f) {
            if (t == nullptr)
                return;
            ofstream ofs(f, ios::binary);
            t->save_data(ofs);
        };
        // the environments live in partition files loaded into frame 1 when
        // used (the right one's load replaced the left one's): reload each
        // before writing it
        if (me->envs[isite]->left != nullptr) {
            if (isite != 0)
                frame_<double>()->load_data(
                    1, me->get_left_partition_filename(isite));
            save_t(me->envs[isite]->left, d + "/lenv.bin");
        }
        if (me->envs[isite]->right != nullptr) {
            frame_<double>()->load_data(
                1, me->get_right_partition_filename(isite));
            save_t(me->envs[isite]->right, d + "/renv.bin");
        }
        frame_<double>()->activate(0);
        save_t(h->op->lopt, d + "/lblk.bin");
        save_t(h->op->ropt, d + "/rblk.bin");
        if (isite > 0) {
            me->ket->load_tensor(isite - 1);
            me->ket->tensors[isite - 1]->save_data(d + "/mps_left.bin", true);
            me->ket->unload_tensor(isite - 1);
        }
        if (isite + 2 < me->n_sites) {
            me->ket->load_tensor(isite + 2);
            me->ket->tensors[isite + 2]->save_data(d + "/mps_right.bin",
                                                   true);
            me->ket->unload_tensor(isite + 2);
        }
        ofstream m(d + "/env.txt");
        m << isite << " " << (me->envs[isite]->left != nullptr) << " "
          << (me->envs[isite]->right != nullptr) << "\n";
    }
    // Density-matrix truncation of the Davidson vector by the reference
    // (MovingEnvironment::density_matrix with noise 0 and
    // split_density_matrix with normalize, as the sweep calls them,
    // src/dmrg/sweep_algorithm.hpp:430-433), in three cases:
    //   sR: trace_right (forward), Physical, k = 2/3 of all states;
    //   sL: trace_left (backward), Physical, k = 1/2;
    //   sK: trace_left, Reduced | KeepOne * 1, k = 1/3;
    //   sW: trace_right, wavefunction noise 1e-6, k = 2/3;
    //   sD: trace_left, density-matrix noise 1e-6, k = 2/3;
    //   sE / sF: trace_right / trace_left, density-matrix noise 1e-6,
    //       every state kept (k = all).
    // Written per case: the density matrix (info + data, before
    // diagonalisation), the two factors (info + data), the error, the
    // wavefunction spectra and the settings.
    void dump_split(const shared_ptr<SparseMatrix<S, double>> &wfn,
                    const string &d) const {
        struct Case {
            const char *tag;
            bool trace_right;
            int num, den;
            TruncationTypes tt;
            double noise;
            NoiseTypes nt;
        };
        // noisy cases draw from block2's generator freshly seeded with 4321
        const Case cases[7] = {
            {"sR", true, 2, 3, TruncationTypes::Physical, 0.0, NoiseTypes::None},
            {"sL", false, 1, 2, TruncationTypes::Physical, 0.0, NoiseTypes::None},
            {"sK", false, 1, 3,
             TruncationTypes::Reduced | TruncationTypes::KeepOne, 0.0,
             NoiseTypes::None},
            {"sW", true, 2, 3, TruncationTypes::Physical, 1E-6,
             NoiseTypes::Wavefunction},
            {"sD", false, 2, 3, TruncationTypes::Physical, 1E-6,
             NoiseTypes::DensityMatrix},
            // density-matrix noise with every state kept (k = all), forward
            // and backward: the states dropped are the ones the noise
            // leaves below the cutoff
            {"sE", true, 1, 1, TruncationTypes::Physical, 1E-6,
             NoiseTypes::DensityMatrix},
            {"sF", false, 1, 1, TruncationTypes::Physical, 1E-6,
             NoiseTypes::DensityMatrix}};
        const double cutoff = 1E-14;
        for (const Case &c : cases) {
            const string t = string(c.tag) + "_";
            if (c.noise != 0.0)
                Random::rand_seed(4321);
            auto dm = MovingEnvironment<S, double, double>::density_matrix(
                S(0), wfn, c.trace_right, c.noise, c.nt);
            InfoPack<S> dp;
            dp.add(dm->info);
            dp.write(d, t + "dm");
            write_npy(d + "/" + t + "dm_data.npy", dm->data, dm->total_memory);
            int k_total = 0;
            for (int i = 0; i < dm->info->n; i++)
                k_total += dm->info->n_states_bra[i];
            const int k = max(1, k_total * c.num / c.den);
            shared_ptr<SparseMatrix<S, double>> left, right;
            vector<double> spectra;
            const double err =
                MovingEnvironment<S, double, double>::split_density_matrix(
                    dm, wfn, k, c.trace_right, true, left, right, cutoff, true,
                    spectra, c.tt);
            InfoPack<S> lp, rp;
            lp.add(left->info);
            rp.add(right->info);
            lp.write(d, t + "left");
            rp.write(d, t + "right");
            write_npy(d + "/" + t + "left_data.npy", left->data,
                      left->total_memory);
            write_npy(d + "/" + t + "right_data.npy", right->data,
                      right->total_memory);
            write_npy(d + "/" + t + "spectra.npy", spectra);
            vector<double> scal = {err, (double)k, cutoff,
                                   (double)(ubond_t)c.tt,
                                   c.trace_right ? 1.0 : 0.0};
            if (c.noise != 0.0) // noise, its type (block2's NoiseTypes), seed
                scal.insert(scal.end(), {c.noise, (double)(int)c.nt, 4321.0});
            write_npy(d + "/" + t + "settings.npy", scal);
            // the stacks are LIFO: release as the sweep does
            // (src/dmrg/sweep_algorithm.hpp:644-657)
            right->info->deallocate();
            right->deallocate();
            left->info->deallocate();
            left->deallocate();
            dm->info->deallocate();
            dm->deallocate();
        }
    }
    // State-averaged truncation by the reference (density_matrix_with_
    // multi_target + multi_split_density_matrix,
    // src/dmrg/moving_environment.hpp:3582-3608, 4347-4462): two roots of
    // the same symmetry, `a` and `b`
    // (each a one-component SparseMatrixGroup), weights 0.7 / 0.3, forward,
    // Physical, k = 2/3 of all states.  Written with tag sM: density
    // matrix, rotation matrix, each root's new wavefunction, error,
    // spectra, settings.
    void dump_multi_split(const shared_ptr<SparseMatrix<S, double>> &a,
                          const vector<double> &bdata, const string &d,
                          const string &tag = "sM", double noise = 0.0,
                          NoiseTypes nt = NoiseTypes::None) const {
        const string t = tag + "_";
        vector<shared_ptr<SparseMatrixGroup<S, double>>> wfns;
        for (int r = 0; r < 2; r++) {
            auto g = make_shared<SparseMatrixGroup<S, double>>();
            g->allocate(vector<shared_ptr<SparseMatrixInfo<S>>>{a->info});
            const double *src = r == 0 ? a->data : bdata.data();
            memcpy(g->data, src, sizeof(double) * a->total_memory);
            wfns.push_back(g);
        }
        const vector<double> weights = {0.7, 0.3};
        if (noise != 0.0)
            Random::rand_seed(4321);
        auto dm = MovingEnvironment<S, double, double>::
            density_matrix_with_multi_target(S(0), wfns, weights, true, noise,
                                             nt);
        InfoPack<S> dp;
        dp.add(dm->info);
        dp.write(d, t + "dm");
        write_npy(d + "/" + t + "dm_data.npy", dm->data, dm->total_memory);
        int k_total = 0;
        for (int i = 0; i < dm->info->n; i++)
            k_total += dm->info->n_states_bra[i];
        const int k = max(1, k_total * 2 / 3);
        const double cutoff = 1E-14;
        vector<shared_ptr<SparseMatrixGroup<S, double>>> new_wfns;
        shared_ptr<SparseMatrix<S, double>> rot;
        vector<double> spectra;
        const double err =
            MovingEnvironment<S, double, double>::multi_split_density_matrix(
                dm, wfns, k, true, true, new_wfns, rot, cutoff, true, spectra,
                TruncationTypes::Physical);
        InfoPack<S> rp, wp;
        rp.add(rot->info);
        rp.write(d, t + "left");
        write_npy(d + "/" + t + "left_data.npy", rot->data, rot->total_memory);
        wp.add(new_wfns[0]->infos[0]);
        wp.write(d, t + "right");
        vector<double> wd;
        for (auto &g : new_wfns)
            wd.insert(wd.end(), g->data, g->data + g->total_memory);
        write_npy(d + "/" + t + "right_data.npy", wd);
        write_npy(d + "/" + t + "spectra.npy", spectra);
        write_npy(d + "/" + t + "roots.npy", bdata);
        vector<double> scal = {err, (double)k, cutoff, 0.0, 1.0};
        if (noise != 0.0) // noise, its type (block2's NoiseTypes), seed
            scal.insert(scal.end(), {noise, (double)(int)nt, 4321.0});
        write_npy(d + "/" + t + "settings.npy", scal);
        // release in reverse order of allocation on each LIFO stack:
        // data (roots, density matrix, rotation, new roots) and infos
        // (density matrix, rotation, new roots' shared info)
        for (int r = (int)new_wfns.size() - 1; r >= 0; r--)
            new_wfns[r]->deallocate();
        rot->deallocate();
        dm->deallocate();
        for (int r = (int)wfns.size() - 1; r >= 0; r--)
            wfns[r]->deallocate();
        new_wfns[0]->infos[0]->deallocate();
        rot->info->deallocate();
        dm->info->deallocate();
    }
    void dump(const shared_ptr<EffectiveHamiltonian<S, double>> &h, string d,
              int isite) const {
        auto op = h->op;
        InfoPack<S> kp, bp, lp, rp;
        kp.add(h->ket->info);
        bp.add(h->bra->info);
        kp.write(d, "ket");
        bp.write(d, "bra");
        for (auto &x : h->left_op_infos)
            lp.add(x.second);
        for (auto &x : h->right_op_infos)
            rp.add(x.second);
        lp.write(d, "linfo");
        rp.write(d, "rinfo");
        // expression
        shared_ptr<OpExpr<S>> expr = op->mat->data[0];
        if (expr->get_type() == OpTypes::ExprRef)
            expr = dynamic_pointer_cast<OpExprRef<S>>(expr)->op;
        vector<shared_ptr<OpProduct<S, double>>> strings;
        if (expr->get_type() == OpTypes::Sum)
            strings = dynamic_pointer_cast<OpSum<S, double>>(expr)->strings;
        else if (expr->get_type() == OpTypes::Prod)
            strings.push_back(dynamic_pointer_cast<OpProduct<S, double>>(expr));
        else
            assert(false);
        OpSet<S> lset, rset, dlset, drset;
        vector<uint8_t> tconj, tkind, tdconj;
        vector<int32_t> ta, tb, tda, tdb;
        vector<double> tf;
        const bool dleft =
            op->lopt->get_type() == OperatorTensorTypes::Delayed;
        const bool dright =
            op->ropt->get_type() == OperatorTensorTypes::Delayed;
        shared_ptr<DelayedOperatorTensor<S, double>> dopt =
            dleft ? dynamic_pointer_cast<DelayedOperatorTensor<S, double>>(
                        op->lopt)
                  : dright
                        ? dynamic_pointer_cast<
                              DelayedOperatorTensor<S, double>>(op->ropt)
                        : nullptr;
        {
            int nprod = 0, nsp = 0;
            for (auto &s : strings)
                (s->get_type() == OpTypes::SumProd ? nsp : nprod)++;
            cerr << "site " << isite << ": " << nprod << " OpProduct, " << nsp
                 << " OpSumProd strings; lopt "
                 << (op->lopt->get_type() == OperatorTensorTypes::Delayed
                         ? "delayed"
                         : "normal")
                 << ", ropt "
                 << (op->ropt->get_type() == OperatorTensorTypes::Delayed
                         ? "delayed"
                         : "normal")
                 << endl;
        }
        for (auto &s : strings) {
            tconj.push_back(s->conj);
            const bool prod = s->get_type() == OpTypes::Prod;
            ta.push_back(lset.get(op->lopt, s->a, prod ? h->tf : nullptr));
            tb.push_back(rset.get(op->ropt, s->b, prod ? h->tf : nullptr));
            tf.push_back(s->factor);
            if (s->get_type() == OpTypes::SumProd) {
                // delayed three-tensor string (tensor_functions.hpp SumProd
                // dispatch): sub-operators of the delayed tensor
                auto sp = dynamic_pointer_cast<OpSumProd<S, double>>(s);
                assert(dopt != nullptr && sp->ops.size() == 2);
                tkind.push_back(1);
                tda.push_back(dlset.get(dopt->lopt, sp->ops[0]));
                tdb.push_back(drset.get(dopt->ropt, sp->ops[1]));
                tdconj.push_back((uint8_t)((int)sp->conjs[0] |
                                           ((int)sp->conjs[1] << 1)));
            } else {
                assert(s->get_type() == OpTypes::Prod);
                tkind.push_back(0);
                tda.push_back(-1);
                tdb.push_back(-1);
                tdconj.push_back(0);
            }
        }
        write_npy(d + "/term_conj.npy", tconj);
        write_npy(d + "/term_a.npy", ta);
        write_npy(d + "/term_b.npy", tb);
        write_npy(d + "/term_factor.npy", tf);
        write_npy(d + "/term_kind.npy", tkind);
        write_npy(d + "/term_da.npy", tda);
        write_npy(d + "/term_db.npy", tdb);
        write_npy(d + "/term_dconj.npy", tdconj);
        lset.write(d, "l", h->left_op_infos);
        rset.write(d, "r", h->right_op_infos);
        if (dopt != nullptr) {
            dlset.write_own(d, "dl");
            drset.write_own(d, "dr");
            // tensor-product connection lists of the delayed side's infos
            const auto &side = dleft ? h->left_op_infos : h->right_op_infos;
            vector<int32_t> tn, tnc;
            vector<uint64_t> tq, tstride;
            vector<uint32_t> tidx, tia, tib, tic;
            vector<double> tfac;
            for (auto &x : side) {
                auto ci = x.second->cinfo;
                int n5[5] = {-1, -1, -1, -1, -1}, nc = 0;
                if (ci != nullptr) {
                    for (int k = 0; k < 5; k++)
                        n5[k] = ci->n[k];
                    nc = (int)ci->nc;
                    for (int k = 0; k < ci->n[4]; k++) {
                        tq.push_back(qdata(ci->quanta[k]));
                        tidx.push_back(ci->idx[k]);
                    }
                    for (int k = 0; k < nc; k++) {
                        tia.push_back(ci->ia[k]);
                        tib.push_back(ci->ib[k]);
                        tic.push_back(ci->ic[k]);
                        tstride.push_back(ci->stride[k]);
                        tfac.push_back(ci->factor[k]);
                    }
                }
                tn.insert(tn.end(), n5, n5 + 5);
                tnc.push_back(nc);
            }
            write_npy(d + "/tp_n.npy", tn);
            write_npy(d + "/tp_nc.npy", tnc);
            write_npy(d + "/tp_quanta.npy", tq);
            write_npy(d + "/tp_idx.npy", tidx);
            write_npy(d + "/tp_ia.npy", tia);
            write_npy(d + "/tp_ib.npy", tib);
            write_npy(d + "/tp_ic.npy", tic);
            write_npy(d + "/tp_stride.npy", tstride);
            write_npy(d + "/tp_factor.npy", tfac);
        }
        // reference ConnectionInfo (wfn) and recomputed diag ConnectionInfo
        assert(h->wfn_infos.size() >= 1 && h->wfn_infos[0] != nullptr);
        write_cinfo<S>(d, "cwfn", h->wfn_infos[0]);
        {
            vector<S> msl = Partition<S, double>::get_uniq_labels({h->hop_mat});
            auto msubsl = Partition<S, double>::get_uniq_sub_labels(
                op->mat, h->hop_mat, msl, h->hop_left_vacuum);
            vector<uint8_t> sconj;
            vector<uint64_t> sq;
            for (auto &x : msubsl[0])
                sconj.push_back(x.first), sq.push_back(qdata(x.second));
            write_npy(d + "/subl_conj.npy", sconj);
            write_npy(d + "/subl_q.npy", sq);
            auto di = make_shared<
                typename SparseMatrixInfo<S>::ConnectionInfo>();
            di->initialize_diag(h->ket->info->delta_quantum, h->opdq,
                                msubsl[0], h->left_op_infos,
                                h->right_op_infos, h->diag->info,
                                h->tf->opf->cg);
            write_cinfo<S>(d, "cdiag", di);
            di->deallocate();
        }
        size_t nw = h->ket->total_memory;
        vector<double> psi(h->ket->data, h->ket->data + nw);
        write_npy(d + "/psi.npy", psi);
        write_npy(d + "/diag.npy", h->diag->data, h->diag->total_memory);
        // H*psi by the reference in SeqTypes::None mode (scalar GEMM chain)
        vector<double> hpsi(h->bra->total_memory, 0.0);
        SeqTypes mode = h->tf->opf->seq->mode;
        h->tf->opf->seq->mode = SeqTypes::None;
        h->tf->opf->seq->cumulative_nflop = 0;
        Timer t;
        t.get_time();
        (*h)(GMatrix<double>(psi.data(), (MKL_INT)nw, 1),
             GMatrix<double>(hpsi.data(), (MKL_INT)hpsi.size(), 1));
        double thpsi = t.get_time();
        size_t nflop = h->tf->opf->seq->cumulative_nflop;
        h->tf->opf->seq->cumulative_nflop = 0;
        h->tf->opf->seq->mode = mode;
        write_npy(d + "/hpsi.npy", hpsi);
#ifdef B2G_SHIM_CHECK
        shim_check<S>(h, d, nflop);
#endif
        // a second, random input vector (seeded) and its H*x
        vector<double> x2(nw), hx2(h->bra->total_memory, 0.0);
        {
            mt19937 gen(1234 + isite);
            uniform_real_distribution<double> u(-1.0, 1.0);
            for (auto &v : x2)
                v = u(gen);
            h->tf->opf->seq->mode = SeqTypes::None;
            (*h)(GMatrix<double>(x2.data(), (MKL_INT)nw, 1),
                 GMatrix<double>(hx2.data(), (MKL_INT)hx2.size(), 1));
            h->tf->opf->seq->mode = mode;
        }
        write_npy(d + "/x2.npy", x2);
        write_npy(d + "/hx2.npy", hx2);
        // Davidson on this site (psi restored afterwards)
        auto dres = h->eigs(nullptr, false, G.dump_conv, 0.0, 5000, -1, 2, 50,
                            DavidsonTypes::Normal, 0.0);
        vector<double> psi_dav(h->ket->data, h->ket->data + nw);
        // the Davidson vector in block2's own file format (info + data,
        // SparseMatrix::save_data with save_info)
        h->ket->save_data(d + "/ket_dav.smat", true);
        // the whole effective problem in block2's EFFPROB1 file (what
        // EffectiveHamiltonian::save_eff_problem writes; operator data
        // other than delayed expressions are not part of it)
        EffectiveProblem<S, double>::save_problem_file(
            d + "/eff_problem.smat", h->op, h->cmat, h->vmat, h->opdq,
            h->operator_quanta, h->wfn_infos, isite, dmrg->forward);
        // the environment operator tensors with their numbers
        // (OperatorTensor::save_data), and for a delayed tensor the tensors
        // of its sub-operators
        auto save_opt = [&d](const shared_ptr<OperatorTensor<S, double>> &t,
                             const string &fn) {
            ofstream ofs(d + "/" + fn, ios::binary);
            t->save_data(ofs);
        };
        save_opt(h->op->lopt, "lopt.smat");
        save_opt(h->op->ropt, "ropt.smat");
        for (auto &t : {h->op->lopt, h->op->ropt})
            if (t->get_type() == OperatorTensorTypes::Delayed) {
                auto dt = dynamic_pointer_cast<DelayedOperatorTensor<S, double>>(t);
                save_opt(dt->lopt, "dlopt.smat");
                save_opt(dt->ropt, "dropt.smat");
            }
        dump_split(h->ket, d);
        {
            // second root: the seeded vector x2, normalised
            vector<double> xb(x2);
            double nrm = 0;
            for (double v : xb)
                nrm += v * v;
            for (double &v : xb)
                v /= sqrt(nrm);
            dump_multi_split(h->ket, xb, d);
            dump_multi_split(h->ket, xb, d, "sN", 1E-6,
                             NoiseTypes::Wavefunction);
            dump_multi_split(h->ket, xb, d, "sQ", 1E-6,
                             NoiseTypes::DensityMatrix);
        }
        memcpy(h->ket->data, psi.data(), nw * sizeof(double));
        write_npy(d + "/psi_dav.npy", psi_dav);
        ofstream m(d + "/meta.txt");
        m << setprecision(17);
        m << "sym " << G.sym << "\n";
        m << "site " << isite << "\n";
        m << "sweep " << dmrg->isweep << "\n";
        m << "forward " << (int)dmrg->forward << "\n";
        m << "opdq " << qdata(h->opdq) << "\n";
        m << "cdq " << qdata(h->ket->info->delta_quantum) << "\n";
        m << "vdq " << qdata(h->bra->info->delta_quantum) << "\n";
        m << "hop_left_vacuum " << qdata(h->hop_left_vacuum) << "\n";
        m << "ket_fermion " << (int)h->ket->info->is_fermion << "\n";
        m << "nflop_mac " << nflop << "\n";
        m << "t_hpsi_s " << thpsi << "\n";
        m << "dav_conv " << G.dump_conv << "\n";
        m << "dav_e " << get<0>(dres) << "\n";
        m << "dav_ndav " << get<1>(dres) << "\n";
        m << "dav_nflop_mac " << get<2>(dres) << "\n";
        m << "dav_t_s " << get<3>(dres) << "\n";
        m << "const_e " << dmrg->me->mpo->const_e << "\n";
        m << "threads " << G.threads << "\n";
        m << "delayed_side " << (dleft ? 1 : dright ? 2 : 0) << "\n";
    }
};

template <typename S> static S make_target(int n, int twos, int pg);
template <> SZ make_target<SZ>(int n, int twos, int pg) {
    return SZ(n, twos, pg);
}
template <> SU2 make_target<SU2>(int n, int twos, int pg) {
    return SU2(n, twos, pg);
}

// --save-sweep: the inputs of the sweeps in block2's own file formats
// (MPO::save_data, src/dmrg/mpo.hpp:795-886; MPSInfo::save_data,
// src/dmrg/mps.hpp:169-191; SparseMatrix::save_data with its info,
// src/core/sparse_matrix.hpp:954-989), and the MPS layout (mps.txt: sites,
// center, dot, canonical form).
template <typename S>
static void save_sweep_inputs(const shared_ptr<MPO<S, double>> &mpo,
                              const shared_ptr<MPS<S, double>> &mps,
                              const shared_ptr<MPSInfo<S>> &info) {
    string d = G.outdir + "/sweep";
    mkdir(d.c_str(), 0755);
    mpo->save_data(d + "/mpo.bin");
    info->save_data(d + "/mps_info.bin");
    for (int i = 0; i < mps->n_sites; i++)
        if (mps->tensors[i] != nullptr)
            mps->tensors[i]->save_data(d + "/mps_" + to_string(i) + ".bin",
                                       true);
    ofstream m(d + "/mps.txt");
    m << mps->n_sites << " " << mps->center << " " << mps->dot << " "
      << mps->canonical_form << "\n";
}

template <typename S> static int run() {
    shared_ptr<FCIDUMP<double>> fcidump = make_shared<FCIDUMP<double>>();
    fcidump->read(G.fcidump);
    vector<uint8_t> orbsym = fcidump->template orb_sym<uint8_t>();
    PGTypes pg = PGTypes::D2H;
    transform(orbsym.begin(), orbsym.end(), orbsym.begin(),
              [pg](uint8_t x) { return (uint8_t)PointGroup::swap_pg(pg)(x); });
    S vacuum(0);
    int norb = fcidump->n_sites();
    S target = make_target<S>(fcidump->n_elec(), fcidump->twos(),
                              PointGroup::swap_pg(pg)(fcidump->isym()));
    auto hamil =
        make_shared<HamiltonianQC<S, double>>(vacuum, norb, orbsym, fcidump);
    Timer t;
    t.get_time();
    shared_ptr<MPO<S, double>> mpo = make_shared<MPOQC<S, double>>(
        hamil, QCTypes::Conventional, "HQC", norb / 2 / 2 * 2);
    mpo->basis = hamil->basis;
    mpo = make_shared<SimplifiedMPO<S, double>>(
        mpo, make_shared<RuleQC<S, double>>(), true, true,
        OpNamesSet({OpNames::R, OpNames::RD}));
    cerr << "MPO done T = " << t.get_time() << endl;
    Random::rand_seed(G.seed);
    auto mps_info =
        make_shared<MPSInfo<S>>(mpo->n_sites, vacuum, target, mpo->basis);
    mps_info->set_bond_dimension(G.bdims[0]);
    auto mps = make_shared<MPS<S, double>>(mpo->n_sites, 0, 2);
    mps->initialize(mps_info);
    mps->random_canonicalize();
    if (G.save_sweep)
        save_sweep_inputs<S>(mpo, mps, mps_info);
    mps->save_mutable();
    mps->deallocate();
    mps_info->save_mutable();
    mps_info->deallocate_mutable();
    auto me =
        make_shared<MovingEnvironment<S, double, double>>(mpo, mps, mps, "DMRG");
    me->init_environments(false);
    if (G.delayed)
        me->delayed_contraction = OpNamesSet::normal_ops();
    me->cached_contraction = true;
    vector<ubond_t> bdims(G.bdims.begin(), G.bdims.end());
    vector<double> noises(G.noises.begin(), G.noises.end());
    auto dmrg = make_shared<DMRG<S, double, double>>(me, bdims, noises);
    dmrg->iprint = 2;
    dmrg->noise_type = G.noise_type == "pert"
                           ? NoiseTypes::ReducedPerturbativeCollected
                           : NoiseTypes::DensityMatrix;
    dmrg->decomp_type = DecompositionTypes::DensityMatrix;
    dmrg->davidson_soft_max_iter = 4000;
    auto dumper = make_shared<Dumper<S>>();
    dumper->dmrg = dmrg.get();
    callback_() = dumper;
    if (G.noise_seed != 0)
        Random::rand_seed(G.noise_seed);
    if (G.dav_thrd > 0)
        dmrg->davidson_conv_thrds = vector<double>(G.nsweeps, G.dav_thrd);
    double energy = dmrg->solve(G.nsweeps, mps->center == 0, G.dav_conv * 0.1);
    callback_() = make_shared<CallbackKernel>();
    ofstream ofs(G.outdir + "/energies.txt");
    ofs << setprecision(17);
    ofs << "final " << energy << "\n";
    for (size_t i = 0; i < dmrg->sweep_energies.size(); i++)
        if (dmrg->sweep_energies[i].size() != 0)
            ofs << "sweep " << i << " " << dmrg->sweep_energies[i][0] << " "
                << (i < dmrg->sweep_time.size() ? dmrg->sweep_time[i] : 0.0)
                << "\n";
    ofs << "dumped " << dumper->ndumped << "\n";
    ofs << "sweep_s";
    for (double t : dumper->sweep_s)
        ofs << " " << t;
    ofs << "\nthreads " << G.threads << "\n";
    if (G.save_sweep) {
        write_npy(G.outdir + "/sweep/steps.npy", dumper->steps);
        ofstream sf(G.outdir + "/sweep/settings.txt");
        sf << setprecision(17) << "tol " << G.dav_conv * 0.1 << "\n"
           << "nsweeps " << G.nsweeps << "\n"
           << "noise_type " << G.noise_type << "\n"
           << "noise_seed " << G.noise_seed << "\n"
           << "final " << energy << "\n";
        sf << "bdims";
        for (int b : G.bdims)
            sf << " " << b;
        sf << "\nnoises";
        for (double x : G.noises)
            sf << " " << x;
        sf << "\ndav_conv";
        for (double x : dmrg->davidson_conv_thrds)
            sf << " " << x;
        sf << "\n";
    }
    cerr << "E = " << setprecision(15) << energy << endl;
    mps_info->deallocate();
    me->remove_partition_files();
    mpo->deallocate();
    hamil->deallocate();
    fcidump->deallocate();
    return 0;
}

// --load-problem DIR: an effective Hamiltonian handed over as arrays (the
// layout of the dumps above: ket_/bra_/linfo_/rinfo_ infos, lop_/rop_
// operators, term_ strings, scalars.npy = symmetry, H's delta quantum,
// hop_left_vacuum; x.npy = n_rhs stacked input vectors) is built into
// block2's own objects (SparseMatrixInfo, SparseMatrix, OpElement keys,
// OpSum of OpProduct, DelayedOperatorTensor) and block2's
// EffectiveHamiltonian (src/dmrg/effective_hamiltonian.hpp:126-224) is
// constructed on it.  Written back: diag.npy (its diagonal), hx.npy (H x for
// each input by block2's Tasked replay, the path eigs takes,
// effective_hamiltonian.hpp:514-521 -> batch_gemm.hpp:1570-1691), and
// timing.txt (threads, block2's multiply-add count of one product, the
// duration of each of `reps` timed products of the first vector, the
// constructor's duration).
template <typename S>
static shared_ptr<SparseMatrixInfo<S>>
load_info(const string &d, const string &p, int i, bool wfn) {
    static map<string, vector<uint64_t>> c64;
    static map<string, vector<uint32_t>> c32;
    auto g64 = [&](const string &k) -> const vector<uint64_t> & {
        string f = d + "/" + p + "_" + k + ".npy";
        if (!c64.count(f))
            c64[f] = read_npy<uint64_t>(f);
        return c64[f];
    };
    auto g32 = [&](const string &k) -> const vector<uint32_t> & {
        string f = d + "/" + p + "_" + k + ".npy";
        if (!c32.count(f))
            c32[f] = read_npy<uint32_t>(f);
        return c32[f];
    };
    const uint32_t n = g32("n")[i], off = g32("off")[i];
    auto info = make_shared<SparseMatrixInfo<S>>();
    info->allocate((int)n);
    info->delta_quantum.data = g64("dq")[i];
    info->is_fermion = g32("fermion")[i] != 0;
    info->is_wavefunction = wfn;
    for (uint32_t j = 0; j < n; j++) {
        info->quanta[j].data = g64("quanta")[off + j];
        info->n_states_bra[j] = (ubond_t)g32("nbra")[off + j];
        info->n_states_ket[j] = (ubond_t)g32("nket")[off + j];
    }
    // block2's own offsets must be the ones the arrays were laid out with
    const uint64_t total = info->template get_total_memory<double>();
    for (uint32_t j = 0; j < n; j++)
        if (info->block_shifts[j] != g32("shift")[off + j])
            throw runtime_error("block shifts differ from block2's in " + p);
    (void)total;
    return info;
}

template <typename S> static int run_load() {
    const string d = G.load_dir;
    vector<uint64_t> sc = read_npy<uint64_t>(d + "/scalars.npy");
    S opdq, hlv;
    opdq.data = sc[1];
    hlv.data = sc[2];
    auto cg = make_shared<CG<S>>();
    auto opf = make_shared<OperatorFunctions<S, double>>(cg);
    auto tf = make_shared<TensorFunctions<S, double>>(opf);
    auto ket_info = load_info<S>(d, "ket", 0, true);
    auto bra_info = load_info<S>(d, "bra", 0, true);
    auto side = [&](const string &ip, const string &op, int half,
                    vector<pair<S, shared_ptr<SparseMatrixInfo<S>>>> &infos,
                    shared_ptr<OperatorTensor<S, double>> &opt,
                    vector<shared_ptr<OpExpr<S>>> &keys,
                    shared_ptr<vector<double>> &data) {
        const vector<uint64_t> idq = read_npy<uint64_t>(d + "/" + ip + "_dq.npy");
        for (size_t i = 0; i < idq.size(); i++) {
            auto info = load_info<S>(d, ip, (int)i, false);
            infos.push_back(make_pair(info->delta_quantum, info));
        }
        const vector<uint64_t> odq = read_npy<uint64_t>(d + "/" + op + "op_dq.npy");
        const vector<double> ofac = read_npy<double>(d + "/" + op + "op_factor.npy");
        const vector<int32_t> oinfo = read_npy<int32_t>(d + "/" + op + "op_info.npy");
        const vector<uint64_t> ooff = read_npy<uint64_t>(d + "/" + op + "op_off.npy");
        data = make_shared<vector<double>>(read_npy<double>(d + "/" + op + "op_data.npy"));
        opt = make_shared<OperatorTensor<S, double>>();
        for (size_t k = 0; k < odq.size(); k++) {
            S q;
            q.data = odq[k];
            // a distinct name per operator: X with a two-index site label
            auto key = make_shared<OpElement<S, double>>(
                OpNames::X,
                SiteIndex((uint16_t)((k & 0xFFF) | (half << 11)),
                          (uint16_t)(k >> 11)),
                q, 1.0);
            auto m = make_shared<SparseMatrix<S, double>>();
            m->info = infos[oinfo[k]].second;
            m->total_memory = m->info->template get_total_memory<double>();
            if (ooff[k] + m->total_memory > data->size())
                throw runtime_error("operator data outside the buffer");
            m->data = data->data() + ooff[k];
            m->factor = ofac[k];
            shared_ptr<OpExpr<S>> ekey = key;
            opt->ops[ekey] = m;
            keys.push_back(ekey);
        }
    };
    vector<pair<S, shared_ptr<SparseMatrixInfo<S>>>> linfos, rinfos;
    shared_ptr<OperatorTensor<S, double>> lopt, ropt;
    vector<shared_ptr<OpExpr<S>>> lkeys, rkeys;
    shared_ptr<vector<double>> ldata, rdata;
    side("linfo", "l", 0, linfos, lopt, lkeys, ldata);
    side("rinfo", "r", 1, rinfos, ropt, rkeys, rdata);
    const vector<uint8_t> tconj = read_npy<uint8_t>(d + "/term_conj.npy");
    const vector<int32_t> ta = read_npy<int32_t>(d + "/term_a.npy");
    const vector<int32_t> tb = read_npy<int32_t>(d + "/term_b.npy");
    const vector<double> tfac = read_npy<double>(d + "/term_factor.npy");
    vector<shared_ptr<OpProduct<S, double>>> strings;
    for (size_t t = 0; t < ta.size(); t++)
        strings.push_back(make_shared<OpProduct<S, double>>(
            dynamic_pointer_cast<OpElement<S, double>>(lkeys[ta[t]]),
            dynamic_pointer_cast<OpElement<S, double>>(rkeys[tb[t]]),
            tfac[t], tconj[t]));
    auto hop = make_shared<OpElement<S, double>>(OpNames::H, SiteIndex(),
                                                 opdq, 1.0);
    auto hop_mat = make_shared<SymbolicColumnVector<S>>(1);
    (*hop_mat)[0] = hop;
    auto op = make_shared<DelayedOperatorTensor<S, double>>();
    op->dops.push_back(hop);
    op->mat = make_shared<SymbolicColumnVector<S>>(1);
    op->mat->data[0] = make_shared<OpSum<S, double>>(strings);
    op->lopt = lopt;
    op->ropt = ropt;
    const vector<double> xs = read_npy<double>(d + "/x.npy");
    auto ket = make_shared<SparseMatrix<S, double>>();
    ket->allocate(ket_info);
    auto bra = make_shared<SparseMatrix<S, double>>();
    bra->allocate(bra_info);
    const size_t nk = ket->total_memory, nb = bra->total_memory;
    if (xs.size() % nk != 0)
        throw runtime_error("x.npy is not a stack of ket vectors");
    const size_t nrhs = xs.size() / nk;
    Timer t;
    t.get_time();
    auto h = make_shared<EffectiveHamiltonian<S, double>>(
        linfos, rinfos, op, bra, ket, hop, hop_mat, hlv, tf, true);
    const double t_ctor = t.get_time();
    write_npy(d + "/diag.npy", h->diag->data, h->diag->total_memory);
    h->tf->opf->seq->cumulative_nflop = 0;
    h->precompute();
    vector<double> hx(nrhs * nb, 0.0), tt;
    size_t nflop = 0;
    for (size_t r = 0; r < nrhs; r++) {
        h->tf->opf->seq->cumulative_nflop = 0;
        h->tf->operator()(
            GMatrix<double>((double *)xs.data() + r * nk, (MKL_INT)nk, 1),
            GMatrix<double>(hx.data() + r * nb, (MKL_INT)nb, 1), 1.0);
        nflop = h->tf->opf->seq->cumulative_nflop;
    }
    vector<double> y(nb);
    for (int rep = 0; rep < G.reps; rep++) {
        memset(y.data(), 0, nb * sizeof(double));
        t.get_time();
        h->tf->operator()(
            GMatrix<double>((double *)xs.data(), (MKL_INT)nk, 1),
            GMatrix<double>(y.data(), (MKL_INT)nb, 1), 1.0);
        tt.push_back(t.get_time());
    }
    h->post_precompute();
    write_npy(d + "/hx.npy", hx);
    ofstream m(d + "/timing.txt");
    m << setprecision(17) << "threads " << G.threads << "\n"
      << "nflop_mac " << nflop << "\n"
      << "t_ctor_s " << t_ctor << "\n"
      << "t_hpsi_s";
    for (double x : tt)
        m << " " << x;
    m << "\n";
    h->deallocate();
    bra->deallocate();
    ket->deallocate();
    return 0;
}

static vector<double> parse_dlist(const string &s) {
    vector<double> r;
    stringstream ss(s);
    string x;
    while (getline(ss, x, ','))
        r.push_back(stod(x));
    return r;
}

int main(int argc, char **argv) {
    for (int i = 1; i + 1 < argc; i += 2) {
        string k = argv[i], v = argv[i + 1];
        if (k == "--fcidump")
            G.fcidump = v;
        else if (k == "--sym")
            G.sym = v;
        else if (k == "--out")
            G.outdir = v;
        else if (k == "--bdims") {
            G.bdims.clear();
            for (double x : parse_dlist(v))
                G.bdims.push_back((int)x);
        } else if (k == "--noises")
            G.noises = parse_dlist(v);
        else if (k == "--nsweeps")
            G.nsweeps = stoi(v);
        else if (k == "--threads")
            G.threads = stoi(v);
        else if (k == "--dump-sweep")
            G.dump_sweep = stoi(v);
        else if (k == "--dump-sites") {
            for (double x : parse_dlist(v))
                G.dump_sites.push_back((int)x);
        } else if (k == "--dav-conv")
            G.dav_conv = stod(v);
        else if (k == "--dump-conv")
            G.dump_conv = stod(v);
        else if (k == "--noise-type")
            G.noise_type = v;
        else if (k == "--delayed")
            G.delayed = stoi(v) != 0;
        else if (k == "--load-problem")
            G.load_dir = v;
        else if (k == "--reps")
            G.reps = stoi(v);
        else if (k == "--save-sweep")
            G.save_sweep = stoi(v) != 0;
        else if (k == "--seed")
            G.seed = (unsigned)stoul(v);
        else if (k == "--noise-seed")
            G.noise_seed = (unsigned)stoul(v);
        else if (k == "--dav-thrd")
            G.dav_thrd = stod(v);
        else {
            cerr << "unknown option " << k << endl;
            return 2;
        }
    }
    if (G.load_dir != "")
        G.outdir = G.load_dir;
    mkdir(G.outdir.c_str(), 0755);
    frame_<double>() = make_shared<DataFrame<double>>(
        (size_t)1 << 26, (size_t)1 << 33, G.outdir + "/scratch");
    frame_<double>()->use_main_stack = false;
    frame_<double>()->minimal_disk_usage = true;
    threading_() = make_shared<Threading>(
        ThreadingTypes::OperatorBatchedGEMM | ThreadingTypes::Global,
        G.threads, G.threads, 1);
    threading_()->seq_type = SeqTypes::Tasked;
    int r = G.load_dir != ""
                ? (G.sym == "sz" ? run_load<SZ>() : run_load<SU2>())
                : (G.sym == "sz" ? run<SZ>() : run<SU2>());
    frame_<double>()->activate(0);
    frame_<double>() = nullptr;
    return r;
}
