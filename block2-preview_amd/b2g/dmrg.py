"""Two-site DMRG sweeps with every numerical step on the GPU.

The sweep follows block2's DMRG::solve / sweep / update_two_dot
(src/dmrg/sweep_algorithm.hpp:3037-3130, 2556-2705, 814-1182) over
MovingEnvironment (src/dmrg/moving_environment.hpp):

  environment blocking   left_contract / right_contract (:1745-2068)
                         -> b2g_contract_plan (window sums on the device)
  renormalisation        left/right_contract_rotate (:226-460, 460-640)
                         -> b2g_rotate_plan (grouped-GEMM replay)
  normal/complementary   numerical_transform at the MPO's schemer sites
                         -> b2g_transform_plan
  effective Hamiltonian  eff_ham (:2071-2209) -> EffectiveHamiltonian
  Davidson               two_dot_eigs_and_perturb (:1183-1266) -> eigs
  truncation             density_matrix + split_density_matrix
                         -> truncate.split_density_matrix (device FP64)
  wavefunction moves     contract_two_dot (:3361-3404), propagate_wfn
                         (:4490-4516) -> b2g_contract_wfn_plan,
                         b2g_swap_plan

Inputs are block2's own files: the MPO (MPO::save_data) and the starting
MPS (MPSInfo::save_data and each tensor's SparseMatrix::save_data), read
by b2g.mpo.  The operator expressions, site operators and bases come from
the MPO; the library builds every connection list and block layout.

Supported: real FP64, SZ and SU(2), MPOs whose contraction and H_eff
expressions are sums of two-operator products (MPOQC conventional, with
its schemer), density-matrix decomposition with no noise, wavefunction or
density-matrix noise.  Perturbative noise and delayed contraction are not
built here.
"""
import ctypes as C
import time

import numpy as np
import torch

from . import _lib as L
from . import states as Q
from .effham import EffectiveHamiltonian
from .mpo import load_mpo, load_mps
from .problem import InfoSet, OpSet, Problem
from . import truncate as T

ZERO_T = ("zero",)


def _key(e):
    """Identity of an operator element: name, site indices, quantum (its
    scalar factor belongs to the expression)."""
    if e[0] != "elem":
        raise ValueError("expected an operator element, got %s" % e[0])
    return (e[1], e[2], e[4])


def _unref(e):
    return e[2] if e[0] == "exprref" else e


def _strings(e):
    """The strings of an expression: products and SumProds."""
    e = _unref(e)
    if e[0] == "zero":
        return []
    if e[0] in ("prod", "sumprod"):
        return [e]
    if e[0] == "sum":
        for x in e[1]:
            if x[0] not in ("prod", "sumprod"):
                raise NotImplementedError("expression string of type %s"
                                          % x[0])
        return list(e[1])
    raise NotImplementedError("expression of type %s" % e[0])


def _sinfo(si, keep):
    q = np.ascontiguousarray(si.q, np.uint64)
    n = np.ascontiguousarray(si.n, np.uint32)
    keep += [q, n]
    return L.b2g_state_info(len(q), 0, L.ptr(q, C.c_uint64),
                            L.ptr(n, C.c_uint32))


def _cinfo(lay, keep):
    arrs = [np.ascontiguousarray(lay.quanta, np.uint64),
            np.ascontiguousarray(lay.nbra, np.uint32),
            np.ascontiguousarray(lay.nket, np.uint32),
            np.ascontiguousarray(lay.shift, np.uint64)]
    keep += arrs
    return L.b2g_sparse_info(int(lay.dq), int(lay.fermion), len(lay),
                             L.ptr(arrs[0], C.c_uint64),
                             L.ptr(arrs[1], C.c_uint32),
                             L.ptr(arrs[2], C.c_uint32),
                             L.ptr(arrs[3], C.c_uint64))


def _dev(t):
    return C.c_void_p(t.data_ptr())


def _info_set(lays):
    """An InfoSet of the engine's Problem from block layouts."""
    n = [len(x) for x in lays]
    off = np.concatenate([[0], np.cumsum(n)[:-1]]) if n else []
    cat = lambda k, dt: (np.concatenate([getattr(x, k) for x in lays])
                         .astype(dt) if lays else np.zeros(0, dt))
    return InfoSet([x.dq for x in lays], [x.fermion for x in lays],
                   n, off, cat("quanta", np.uint64),
                   cat("nbra", np.uint32), cat("nket", np.uint32),
                   cat("shift", np.uint64))


def _op_set(t):
    return OpSet(t.dq, t.factor, t.info_of, t.off, None)


class OpTensor:
    """The operators of one block (an OperatorTensor's ops): keys in order,
    one block layout per delta quantum (sorted), device FP64 data."""

    def __init__(self, sym, bra, ket, keys, dqs, device, factors=None):
        self.sym, self.bra, self.ket = sym, bra, ket
        self.keys = list(keys)
        self.index = {k: i for i, k in enumerate(self.keys)}
        self.dq = [int(d) for d in dqs]
        self.info_dq = sorted(set(self.dq))
        self.infos = [Q.initialize(sym, bra, ket, d, Q.q_is_fermion(d))
                      for d in self.info_dq]
        self.factor = (np.ones(len(keys)) if factors is None else
                       np.asarray(factors, np.float64))
        self.info_of = [self.info_dq.index(d) for d in self.dq]
        self.off = []
        t = 0
        for i in self.info_of:
            self.off.append(t)
            t += self.infos[i].total
        self.total = t
        self.data = torch.zeros(max(t, 1), dtype=torch.float64, device=device)

    def layout(self, k):
        return self.infos[self.info_of[k]]

    def c_side(self, keep):
        dq = np.ascontiguousarray(self.info_dq, np.uint64)
        keep.append(dq)
        inf = (L.b2g_sparse_info * max(len(self.infos), 1))(
            *[_cinfo(x, keep) for x in self.infos])
        ops = (L.b2g_operator * max(len(self.keys), 1))(
            *[L.b2g_operator(self.dq[k], float(self.factor[k]),
                             self.info_of[k], 0, self.off[k])
              for k in range(len(self.keys))])
        keep += [inf, ops]
        return (len(self.infos), L.ptr(dq, C.c_uint64), inf, len(self.keys),
                ops)

    def extend(self, keys, dqs):
        """This tensor with zero operators `keys` (quanta `dqs`) appended
        after its data; its own operators keep their offsets."""
        r = object.__new__(OpTensor)
        r.sym, r.bra, r.ket = self.sym, self.bra, self.ket
        r.keys = self.keys + list(keys)
        r.index = {k: i for i, k in enumerate(r.keys)}
        r.dq = self.dq + [int(d) for d in dqs]
        new = sorted(set(r.dq) - set(self.info_dq))
        r.info_dq = sorted(set(self.info_dq) | set(new))
        lay = {d: x for d, x in zip(self.info_dq, self.infos)}
        for d in new:
            lay[d] = Q.initialize(self.sym, self.bra, self.ket, d,
                                  Q.q_is_fermion(d))
        r.infos = [lay[d] for d in r.info_dq]
        r.info_of = [r.info_dq.index(d) for d in r.dq]
        r.factor = np.concatenate([self.factor, np.ones(len(keys))])
        r.off = list(self.off)
        t = self.total
        for k in range(len(self.keys), len(r.keys)):
            r.off.append(t)
            t += r.infos[r.info_of[k]].total
        r.total = t
        r.data = torch.zeros(max(t, 1), dtype=torch.float64,
                             device=self.data.device)
        if self.total:
            r.data[:self.total] = self.data[:self.total]
        return r

    def signature(self):
        """The structure a plan depends on: operators, layouts, offsets,
        factors and bases."""
        return (tuple(self.keys), tuple(self.dq), tuple(self.off),
                np.asarray(self.factor, np.float64).tobytes(),
                self.bra.q.tobytes(), self.bra.n.tobytes(),
                self.ket.q.tobytes(), self.ket.n.tobytes(), self.total,
                self.data.numel())

    def structure(self):
        """This tensor's structure without its data (a template)."""
        r = object.__new__(OpTensor)
        r.__dict__.update(self.__dict__)
        r._device = self.data.device
        r.data = None
        return r

    def empty_like(self):
        """A tensor of this structure with zero data (of a fresh tensor's
        length)."""
        r = object.__new__(OpTensor)
        r.__dict__.update(self.__dict__)
        dev = self.data.device if self.data is not None else self._device
        r.data = torch.zeros(max(self.total, 1), dtype=torch.float64,
                             device=dev)
        return r

    def subset(self, keys):
        """The operators `keys` of this tensor (same data, new index)."""
        r = object.__new__(OpTensor)
        r.sym, r.bra, r.ket = self.sym, self.bra, self.ket
        r.keys = list(keys)
        r.index = {k: i for i, k in enumerate(r.keys)}
        pos = [self.index[k] for k in r.keys]
        r.dq = [self.dq[p] for p in pos]
        r.info_dq, r.infos = self.info_dq, self.infos
        r.info_of = [self.info_of[p] for p in pos]
        r.off = [self.off[p] for p in pos]
        r.factor = np.array([self.factor[p] for p in pos])
        r.total, r.data = self.total, self.data
        return r


class SiteOps:
    """The site operators of one MPO tensor on the device."""

    def __init__(self, sym, tensor, site_op_infos, device):
        self.sym = sym
        self.keys, self.dq, self.factor, self.off = [], [], [], []
        lay = {}
        for dq, d in site_op_infos:
            lay[int(dq)] = Q.Layout.from_dict(d, wfn=False)
        self.info_dq = sorted(lay)
        self.infos = [lay[d] for d in self.info_dq]
        chunks, t = [], 0
        for k, m in tensor["ops"]:
            key = _key(k)
            d = int(m["info"]["dq"])
            if not Q.Layout.from_dict(m["info"], wfn=False).same(lay[d]):
                raise ValueError("site operator layout differs from its "
                                 "site_op_infos entry")
            self.keys.append(key)
            self.dq.append(d)
            self.factor.append(float(m["factor"]))
            self.off.append(t)
            data = np.asarray(m["data"], np.float64)
            if data.size != lay[d].total:
                data = np.zeros(lay[d].total)
            chunks.append(data)
            t += data.size
        self.index = {k: i for i, k in enumerate(self.keys)}
        self.info_of = [self.info_dq.index(d) for d in self.dq]
        self.host = np.concatenate(chunks) if chunks else np.zeros(1)
        self.data = torch.tensor(self.host if self.host.size else [0.0],
                                 dtype=torch.float64, device=device)
        self.lmat, self.rmat = tensor["lmat"], tensor["rmat"]

    def c_side(self, keep):
        dq = np.ascontiguousarray(self.info_dq, np.uint64)
        keep.append(dq)
        inf = (L.b2g_sparse_info * max(len(self.infos), 1))(
            *[_cinfo(x, keep) for x in self.infos])
        ops = (L.b2g_operator * max(len(self.keys), 1))(
            *[L.b2g_operator(self.dq[k], self.factor[k], self.info_of[k], 0,
                             self.off[k]) for k in range(len(self.keys))])
        keep += [inf, ops]
        return (len(self.infos), L.ptr(dq, C.c_uint64), inf, len(self.keys),
                ops)

    def block_of(self, k, q):
        """host copy of site operator k's block with composite quanta q"""
        lay = self.infos[self.info_of[k]]
        i = lay.find(q)
        if i < 0:
            return None
        o = self.off[k] + int(lay.shift[i])
        return self.host[o:o + int(lay.nbra[i]) * int(lay.nket[i])]


class MPSTensor:
    """An MPS tensor on the device: block layout and data."""

    def __init__(self, lay, data):
        self.lay, self.data = lay, data


def _layout_of_split(bi, wfn):
    return Q.Layout(bi.quanta, bi.nbra, bi.nket, bi.dq, bi.fermion, wfn)


class DMRG:
    """block2's DMRG (sweep_algorithm.hpp) on the GPU for one MPS.

    DMRG(ctx, symmetry ("sz" / "su2"), mpo_file, mps_dir, bond_dims,
    noises) reads block2's files;
    solve(n_sweeps, forward, tol) runs the sweeps and returns the final
    energy; sweep_energies / energies / discarded_weights / sweep_time hold
    the per-sweep results and `steps` one record per sweep step (sweep,
    step, energy, discarded weight, Davidson iterations, seconds);
    phase_time the seconds spent per phase of the steps (wavefunction moves,
    blocking, setup, davidson, truncation, renormalisation).

    precision: the Davidson's H*psi, "f64", or "f32" / "bf16x3" (FP32 MFMA
    or split-BF16 products inside the Davidson, then an FP64 refinement of
    the converged vector; Gram-Schmidt, the Ritz problem, residuals and the
    truncation stay FP64)."""

    def __init__(self, ctx, symmetry, mpo_file, mps_dir, bond_dims, noises,
                 ubond=4, noise_type="dm", noise_seed=0, precision="f64"):
        if precision not in ("f64", "f32", "bf16x3"):
            raise ValueError("precision: f64, f32 or bf16x3")
        self.precision = precision
        self.phase_time, self.davidson_macs = {}, 0
        self._plans = {}
        self.ctx = ctx
        self.dev = ctx.device
        self.lib = L.lib()
        self.mpo = load_mpo(mpo_file, ubond)
        self.mps = load_mps(mps_dir, ubond)
        m, ps = self.mpo, self.mps
        self.n = m["n_sites"]
        if ps["dot"] != 2:
            raise NotImplementedError("two-site MPS only")
        info = ps["info"]
        self.sym = {"sz": Q.SZ, "su2": Q.SU2}[symmetry]
        s = self.sym
        self.target = int(info["target"])
        self.vacuum = int(info["vacuum"])
        self.basis = [Q.StateInfo.from_dict(b) for b in info["basis"]]
        self.ldf = [Q.StateInfo.from_dict(b) for b in info["left_dims_fci"]]
        self.rdf = [Q.StateInfo.from_dict(b) for b in info["right_dims_fci"]]
        self.hq = int(m["op"][4])
        self.const_e = float(m["const_e"])
        self.site = [SiteOps(s, m["tensors"][i], m["site_op_infos"][i],
                             self.dev) for i in range(self.n)]
        self.schemer = m["schemer"]
        self.bond_dims = list(bond_dims)
        self.noises = list(noises)
        # "pert": ReducedPerturbativeCollected, the Python driver's default
        self.perturbative = noise_type == "pert"
        self.noise_type = {"dm": T.NOISE_DENSITY_MATRIX,
                           "wfn": T.NOISE_WAVEFUNCTION,
                           "pert": 0}[noise_type]
        self.rng = T.Rng(noise_seed) if noise_seed else None
        self.davidson_max_iter = 5000
        self.davidson_soft_max_iter = 4000
        self.davidson_def_min, self.davidson_def_max = 2, 50
        self.cutoff = 1e-14
        self.center = ps["center"]
        self._load_tensors()

    # ------------------------------------------------------------ MPS
    def _load_tensors(self):
        s, n = self.sym, self.n
        vac = Q.StateInfo([self.vacuum], [1])
        self.left_dims = [None] * (n + 1)
        self.right_dims = [None] * (n + 1)
        self.left_dims[0] = vac
        self.right_dims[n] = vac
        self.tensors = [None] * n
        ts = self.mps["tensors"]
        c = self.center
        for i in range(n):
            if ts[i] is None:
                continue
            lay = Q.Layout.from_dict(ts[i]["info"])
            data = torch.tensor(ts[i]["data"] * ts[i]["factor"],
                                dtype=torch.float64, device=self.dev)
            self.tensors[i] = MPSTensor(lay, data)
        for i in range(c + 2, n):
            self.right_dims[i] = Q.extract_state_info(
                s, self.tensors[i].lay, False)
        for i in range(0, c):
            self.left_dims[i + 1] = Q.extract_state_info(
                s, self.tensors[i].lay, True)
        # the two-site tensor at the center
        wl = self._wfn_layout(c)
        if not wl.same(self.tensors[c].lay):
            raise ValueError("center tensor layout differs from the fused "
                             "bases of the MPS")

    def _fused_left(self, i):
        return Q.tensor_product(self.sym, self.left_dims[i], self.basis[i],
                                self.ldf[i + 1])

    def _fused_right(self, i):
        return Q.tensor_product(self.sym, self.basis[i], self.right_dims[i + 1],
                                self.rdf[i])

    def _wfn_layout(self, i):
        return Q.initialize(self.sym, self._fused_left(i),
                            self._fused_right(i + 1), self.target, False,
                            True)

    # ------------------------------------------------------ contractions
    def _run_contract(self, A, B, Cs, bra_a, ket_a, bra_b, ket_b, terms,
                      keep_plan=False):
        keep = []
        d = L.b2g_contract()
        d.symmetry = self.sym
        d.bra_a, d.ket_a = _sinfo(bra_a, keep), _sinfo(ket_a, keep)
        d.bra_b, d.ket_b = _sinfo(bra_b, keep), _sinfo(ket_b, keep)
        d.bra, d.ket = _sinfo(Cs.bra, keep), _sinfo(Cs.ket, keep)
        d.n_a_infos, d.a_info_dq, d.a_infos, d.n_a_ops, d.a_ops = \
            A.c_side(keep)
        d.n_b_infos, d.b_info_dq, d.b_infos, d.n_b_ops, d.b_ops = \
            B.c_side(keep)
        d.n_c_infos, d.c_info_dq, d.c_infos, d.n_c_ops, d.c_ops = \
            Cs.c_side(keep)
        tm = (L.b2g_contract_term * max(len(terms), 1))(
            *[L.b2g_contract_term(c, a, b, cj, f)
              for (c, a, b, cj, f) in terms])
        keep.append(tm)
        d.n_terms, d.terms = len(terms), tm
        h = C.c_void_p()
        t = time.perf_counter()
        L.check(self.lib.b2g_contract_plan(self.ctx.h, C.byref(d), 0,
                                           C.byref(h)))
        self._tick("blocking_plan", t)
        if keep_plan:
            self._exec_winplan(h, A.data, B.data, Cs.data, 0, "blocking_run")
            return h
        try:
            self._exec_winplan(h, A.data, B.data, Cs.data, 0, "blocking_run")
        finally:
            self.lib.b2g_winplan_destroy(h)

    def _exec_winplan(self, h, a, b, c, accumulate, phase):
        t = time.perf_counter()
        torch.cuda.current_stream(self.dev).synchronize()
        L.check(self.lib.b2g_winplan_run(h, _dev(a), None if b is None else
                                         _dev(b), _dev(c), accumulate))
        self.ctx.synchronize()
        self._tick(phase, t)

    def _cached(self, slot, sig):
        ent = self._plans.get(slot)
        return ent if ent is not None and ent["sig"] == sig else None

    def _store(self, slot, ent):
        old = self._plans.pop(slot, None)
        if old is not None:
            self._free_entry(old)
        self._plans[slot] = ent

    def _free_entry(self, ent):
        for k in ("contract", "inter"):
            if ent.get(k):
                self.lib.b2g_winplan_destroy(ent[k])
                ent[k] = None

    def close(self):
        """Releases the plans kept across sweeps."""
        for ent in getattr(self, "_plans", {}).values():
            self._free_entry(ent)
        self._plans = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _block_general(self, i, env, left):
        """The contraction of environment and site i (left: env (x) site,
        else site (x) env), its plans kept per side and site while the
        environment's structure repeats (a sweep at a steady bond
        dimension meets the same structures again)."""
        s, m = self.sym, self.mpo
        slot = ("L" if left else "R", i)
        sig = env.signature()
        site = self.site[i]
        ent = self._cached(slot, sig)
        if ent is not None:
            Cs = ent["Cs"].empty_like()
            src = env
            if ent["inter"] is not None:
                src = env.extend(ent["inter_keys"], ent["inter_dqs"])
                self._exec_winplan(ent["inter"], src.data, None, src.data, 1,
                                   "transform_run")
            a, b = (src, site) if left else (site, src)
            self._exec_winplan(ent["contract"], a.data, b.data, Cs.data, 0,
                               "blocking_run")
            return Cs
        fused = self._fused_left(i) if left else self._fused_right(i)
        side = "left" if left else "right"
        names = m[side + "_operator_names"][i]["data"]
        exprs = m[side + "_operator_exprs"][i]["data"]
        live = [x for x in names if x[0] != "zero"]
        Cs = OpTensor(s, fused, fused, [_key(x) for x in live],
                      [x[4] for x in live], self.dev)
        src, terms, inter = self._contraction_terms(env, site, Cs, names,
                                                    exprs, left)
        if left:
            h = self._run_contract(src, site, Cs, src.bra, src.ket,
                                   self.basis[i], self.basis[i], terms,
                                   keep_plan=True)
        else:
            h = self._run_contract(site, src, Cs, self.basis[i],
                                   self.basis[i], src.bra, src.ket, terms,
                                   keep_plan=True)
        ent = dict(sig=sig, Cs=Cs.structure(), contract=h, inter=None,
                   inter_keys=None, inter_dqs=None)
        if inter is not None:
            ent.update(inter=inter[0], inter_keys=inter[1],
                       inter_dqs=inter[2])
        self._store(slot, ent)
        return Cs

    def _vacuum_env(self):
        vac = Q.StateInfo([self.vacuum], [1])
        e = OpTensor(self.sym, vac, vac, [("I",)], [self.vacuum], self.dev)
        e.data.fill_(1.0)
        return e

    def left_block(self, i, env):
        """Operators of (left environment of site i) (x) site i on the fused
        basis (MovingEnvironment::left_contract; site 0: left_assign)."""
        if i != 0:
            return self._block_general(i, env, True)
        s, m = self.sym, self.mpo
        fused = self._fused_left(i)
        names = m["left_operator_names"][i]["data"]
        live = [x for x in names if x[0] != "zero"]
        Cs = OpTensor(s, fused, fused, [_key(x) for x in live],
                      [x[4] for x in live], self.dev)
        site = self.site[i]
        vac = Q.StateInfo([self.vacuum], [1])
        if i == 0:
            # left_assign: the site tensor's row vector (tensor_functions.hpp
            # :185-221) as products with the vacuum identity
            A = self._vacuum_env()
            lm = site.lmat["data"]
            terms = []
            for x in lm:
                if x[0] == "zero":
                    continue
                terms.append((Cs.index[_key(x)], 0, site.index[_key(x)], 0,
                              1.0))
            self._run_contract(A, site, Cs, vac, vac, self.basis[i],
                               self.basis[i], terms)
            return Cs
        return self._block_general(i, env, True)

    def _contraction_terms(self, env, site, Cs, names, exprs, left):
        """Window-sum terms (c, a, b, conj, factor) of a left (right)
        contraction, environment (x) site (site (x) environment).  A
        SumProd string multiplies a sum of environment operators: that sum
        is formed first as an intermediate operator of the first summand's
        shape (OperatorFunctions::iadd, as TensorFunctions::tensor_product
        does, src/core/tensor_functions.hpp:2220-2271), appended to the
        environment, which is returned with the terms."""
        rows, inters = [], []
        for nm, e in zip(names, exprs):
            if nm[0] == "zero":
                continue
            c = Cs.index[_key(nm)]
            for x in _strings(e):
                f, cj = x[1], x[2]
                if x[0] == "prod":
                    ek = _key(x[3] if left else x[4])
                    sk = _key(x[4] if left else x[3])
                else:
                    a, b, ops, conjs = x[3], x[4], x[6], x[7]
                    if (a is None) != left or (b is None) == left:
                        raise NotImplementedError(
                            "SumProd summing site operators")
                    sk = _key(b if left else a)
                    ek = ("sum", len(inters))
                    inters.append((ops[0], [(_key(o), int(bool(q)), o[3])
                                            for o, q in zip(ops, conjs)]))
                rows.append((c, ek, site.index[sk], cj, f / nm[3]))
        inter = None
        if inters:
            env, inter = self._intermediates(env, inters)
        terms = [(c, env.index[ek], sb, cj, f) if left else
                 (c, sb, env.index[ek], cj, f)
                 for c, ek, sb, cj, f in rows]
        return env, terms, inter

    def _intermediates(self, env, inters):
        """env with the intermediate operators appended: operator j =
        sum of f op(env operator) over inters[j][1] (conj: transposed), in
        the shape of the element inters[j][0] (its quantum); and the kept
        plan with the appended keys and quanta."""
        keys = [("sum", j) for j in range(len(inters))]
        dqs = [int(o[4]) for o, _ in inters]
        out = env.extend(keys, dqs)
        terms = []
        for j, (_, parts) in enumerate(inters):
            c = out.index[("sum", j)]
            for k, cj, f in parts:
                terms.append((c, env.index[k], cj, f))
        h = self._run_transform(out, terms, keep_plan=True)
        return out, (h, keys, dqs)

    def right_block(self, i, env):
        """Operators of site i (x) (right environment of site i) on the
        fused basis (right_contract; last site: right_assign)."""
        if i != self.n - 1:
            return self._block_general(i, env, False)
        s, m = self.sym, self.mpo
        fused = self._fused_right(i)
        names = m["right_operator_names"][i]["data"]
        live = [x for x in names if x[0] != "zero"]
        Cs = OpTensor(s, fused, fused, [_key(x) for x in live],
                      [x[4] for x in live], self.dev)
        site = self.site[i]
        vac = Q.StateInfo([self.vacuum], [1])
        if i == self.n - 1:
            B = self._vacuum_env()
            rm = site.rmat["data"]
            terms = []
            for x in rm:
                if x[0] == "zero":
                    continue
                terms.append((Cs.index[_key(x)], site.index[_key(x)], 0, 0,
                              1.0))
            self._run_contract(site, B, Cs, self.basis[i], self.basis[i], vac,
                               vac, terms)
            return Cs
        return self._block_general(i, env, False)

    def rotate(self, blk, rot, bond, trans):
        """Renormalised operators on `bond` (left_rotate: trans False,
        rot^T A rot; right_rotate: trans True, rot A rot^T)."""
        s = self.sym
        out = OpTensor(s, bond, bond, blk.keys, blk.dq, self.dev,
                       blk.factor)
        out.factor = np.ones(len(blk.keys))
        keep = []
        na, adq, ainf, nop, aops = blk.c_side(keep)
        nc, cdq, cinf, _, cops = out.c_side(keep)
        rbuf = C.c_void_p()
        L.check(self.lib.b2g_buf_wrap(self.ctx.h, _dev(rot.data),
                                      rot.data.numel(), L.B2G_F64,
                                      C.byref(rbuf)))
        plan = C.c_void_p()
        try:
            ri = _cinfo(rot.lay, keep)
            t = time.perf_counter()
            L.check(self.lib.b2g_rotate_plan(
                self.ctx.h, s, 1 if trans else 0, C.byref(ri), rbuf, na, adq,
                ainf, nc, cdq, cinf, nop, aops, cops,
                blk.total, out.total, 0, C.byref(plan)))
            t = self._tick("rotation_plan", t)
            torch.cuda.current_stream(self.dev).synchronize()
            L.check(self.lib.b2g_hpsi(self.ctx.h, plan, _dev(blk.data),
                                      _dev(out.data), 1, 1.0, 0,
                                      L.PREC["f64"]))
            self.ctx.synchronize()
            self._tick("rotation_run", t)
        finally:
            if plan:
                self.lib.b2g_plan_destroy(plan)
            self.lib.b2g_buf_destroy(rbuf)
        return out

    def transform(self, env, names, exprs):
        """TensorFunctions::numerical_transform + post_numerical_transform:
        the operators `names` of the environment, new ones formed from
        existing ones by `exprs`."""
        s = self.sym
        live = [(x, e) for x, e in zip(names["data"], exprs["data"])
                if x[0] != "zero"]
        keys = [_key(x) for x, _ in live]
        new = [k for k in keys if k not in env.index]
        dqs = {k: x[4] for k, (x, _) in zip(keys, live)}
        if new:
            # an environment holding both sets (the new operators zero)
            env = env.extend(new, [dqs[k] for k in new])
        terms = []
        for j, (nm, e) in enumerate(live):
            e = _unref(e)
            if e[0] == "zero":
                continue
            fn = nm[3]
            if e[0] != "sum":
                raise NotImplementedError("transform expression %s" % e[0])
            for x in e[1]:
                if x[0] != "prod" or x[4] is not None:
                    raise NotImplementedError("transform string")
                terms.append((env.index[keys[j]], env.index[_key(x[3])],
                              x[2] & 1, x[1] / fn))
        self._run_transform(env, terms)
        return env.subset(keys)

    def _run_transform(self, env, terms, keep_plan=False):
        """env.c += f op(env.a) for the terms (c, a, conj, f), in place
        (OperatorFunctions::iadd on the device); keep_plan: returns the
        plan instead of releasing it."""
        if not terms:
            return None
        keep = []
        ni, dq, inf, nop, ops = env.c_side(keep)
        tm = (L.b2g_transform_term * len(terms))(
            *[L.b2g_transform_term(c, a, cj, 0, f)
              for (c, a, cj, f) in terms])
        h = C.c_void_p()
        t = time.perf_counter()
        L.check(self.lib.b2g_transform_plan(
            self.ctx.h, self.sym, ni, dq, inf, nop, ops,
            len(terms), tm, 0, C.byref(h)))
        self._tick("transform_plan", t)
        if keep_plan:
            self._exec_winplan(h, env.data, None, env.data, 1,
                               "transform_run")
            return h
        try:
            self._exec_winplan(h, env.data, None, env.data, 1,
                               "transform_run")
        finally:
            self.lib.b2g_winplan_destroy(h)
        return None

    # ------------------------------------------------- wavefunction moves
    def _contract_wfn(self, lmat, l_is_wfn, rmat, out_lay):
        keep = []
        li, ri, oi = _cinfo(lmat.lay, keep), _cinfo(rmat.lay, keep), \
            _cinfo(out_lay, keep)
        h = C.c_void_p()
        L.check(self.lib.b2g_contract_wfn_plan(
            self.ctx.h, self.sym, C.byref(li), 1 if l_is_wfn else 0,
            C.byref(ri), C.byref(oi), 1.0, C.byref(h)))
        out = torch.zeros(max(out_lay.total, 1), dtype=torch.float64,
                          device=self.dev)
        try:
            torch.cuda.current_stream(self.dev).synchronize()
            L.check(self.lib.b2g_gemmlist_run(h, _dev(lmat.data),
                                              _dev(rmat.data), _dev(out)))
            self.ctx.synchronize()
        finally:
            self.lib.b2g_gemmlist_destroy(h)
        return MPSTensor(out_lay, out)

    def _swap(self, wfn, i, to_left):
        """MPSInfo::swap_wfn_to_fused_left / right at site i."""
        s = self.sym
        l, mm, r = self.left_dims[i], self.basis[i], self.right_dims[i + 1]
        lm = Q.tensor_product(s, l, mm, self.ldf[i + 1])
        mr = Q.tensor_product(s, mm, r, self.rdf[i])
        if to_left:
            new = Q.initialize(s, lm, r, wfn.lay.dq, wfn.lay.fermion, True)
            old_f, new_f = mr, lm
        else:
            new = Q.initialize(s, l, mr, wfn.lay.dq, wfn.lay.fermion, True)
            old_f, new_f = lm, mr
        keep = []
        args = [_sinfo(x, keep) for x in (l, mm, r, old_f, new_f)]
        oi, ni = _cinfo(wfn.lay, keep), _cinfo(new, keep)
        h = C.c_void_p()
        L.check(self.lib.b2g_swap_plan(
            self.ctx.h, s, 1 if to_left else 0, *[C.byref(a) for a in args],
            C.byref(oi), C.byref(ni), 0, C.byref(h)))
        out = torch.zeros(max(new.total, 1), dtype=torch.float64,
                          device=self.dev)
        try:
            torch.cuda.current_stream(self.dev).synchronize()
            L.check(self.lib.b2g_winplan_run(h, _dev(wfn.data), None,
                                             _dev(out), 0))
            self.ctx.synchronize()
        finally:
            self.lib.b2g_winplan_destroy(h)
        return MPSTensor(new, out)

    # ------------------------------------------------------ environments
    def _schemer_left(self, i):
        sc = self.schemer
        return sc is not None and i == sc["left_trans_site"]

    def _schemer_right(self, i):
        sc = self.schemer
        return sc is not None and i == sc["right_trans_site"]

    def init_environments(self):
        """Right environments for a forward sweep from the center
        (MovingEnvironment::init_environments, right_contract_rotate)."""
        n, c = self.n, self.center
        self.lenv = [None] * (n + 1)
        self.renv = [None] * (n + 1)
        for j in range(n - 1, c + 1, -1):
            blk = self.right_block(j, self.renv[j + 1])
            env = self.rotate(blk, self.tensors[j], self.right_dims[j], True)
            if self._schemer_right(j):
                sc = self.schemer
                env = self.transform(env, sc["right_new_operator_names"],
                                     sc["right_new_operator_exprs"])
            self.renv[j] = env
        if c != 0:
            raise NotImplementedError("starting center must be site 0")

    # ------------------------------------------------------------- sweep
    def _eff_ham(self, i, lblk, rblk, lay):
        s, m = self.sym, self.mpo
        strings = _strings(m["middle_operator_exprs"][i]["data"][0])
        fn = m["middle_operator_names"][i]["data"][0][3]
        conj, ta, tb, tf = [], [], [], []
        for x in strings:
            if x[0] != "prod":
                # SumProd strings of the Hamiltonian belong to delayed
                # operator tensors, which this driver does not form
                raise NotImplementedError("Hamiltonian string %s" % x[0])
            f, cj, a, b = x[1], x[2], x[3], x[4]
            conj.append(cj)
            ta.append(lblk.index[_key(a)])
            tb.append(rblk.index[_key(b)])
            tf.append(f / fn)

        iset, oset = _info_set, _op_set
        w = iset([lay])
        pb = Problem(s, self.hq, 0, w, w, iset(lblk.infos),
                     iset(rblk.infos), oset(lblk), oset(rblk), conj, ta, tb,
                     tf, dict(site=i))
        return EffectiveHamiltonian(pb, self.ctx,
                                    fp32=self.precision == "f32",
                                    device_data=(lblk.data, rblk.data))

    def _tick(self, name, t):
        """Adds the seconds since t to phase `name`; returns now."""
        now = time.perf_counter()
        self.phase_time[name] = self.phase_time.get(name, 0.0) + now - t
        return now

    def _perturbative_noise(self, i, lblk, rblk, lay, x, forward, noise):
        """EffectiveHamiltonian::perturbative_noise for ReducedPerturbative
        noise (src/dmrg/effective_hamiltonian.hpp:263-433) and its scaling
        (MovingEnvironment::scale_perturbative_noise and density_matrix,
        src/dmrg/moving_environment.hpp:3554-3577, 3698-3713).

        Forward (the density matrix traces the right side), every string
        f (a (x) b) of the Hamiltonian contributes f op(a) psi with the
        right side's identity in place of b
        (TensorFunctions::tensor_product_partial_multiply,
        src/core/tensor_functions.hpp:366-804, TraceTypes::Right); backward
        the mirror, psi op(b)^T with the left identity.  The contributions
        are summed per quantum q of the result (q = psi's + op(a)'s), one
        perturbed wavefunction per q, each an H*psi of the engine with the
        identity as the other operator.  Each is normalised, all are scaled
        to a total squared norm of `noise`, and all but the one of the
        lowest q are returned, as (layout, data), for the density
        matrix."""
        s, m = self.sym, self.mpo
        ident = (2, 0, self.vacuum)        # OpNames::I, no site, vacuum
        side = rblk if forward else lblk
        if ident not in side.index:
            return []                      # nothing to perturb here
        iid = side.index[ident]
        strings = _strings(m["middle_operator_exprs"][i]["data"][0])
        fn = m["middle_operator_names"][i]["data"][0][3]
        groups = {}
        for st in strings:
            if st[0] != "prod":
                raise NotImplementedError("Hamiltonian string %s" % st[0])
            f, cj, a, b = st[1], st[2], st[3], st[4]
            if forward:
                c1, op = cj & 1, a
                term = (c1, lblk.index[_key(a)], iid)
            else:
                c1, op = (cj >> 1) & 1, b
                term = (c1 << 1, iid, rblk.index[_key(b)])
            opdq = Q.q_neg(s, int(op[4])) if c1 else int(op[4])
            groups.setdefault((c1, opdq), []).append(term + (f / fn,))
        vdqs = {}
        for (c1, opdq) in groups:
            pk = Q.q_add(s, int(lay.dq), opdq)
            vdqs[(c1, opdq)] = [Q.q_at(s, pk, k)
                                for k in range(Q.q_count(s, pk))]
        labels = sorted(set(v for vs in vdqs.values() for v in vs))
        fl, fr = self._fused_left(i), self._fused_right(i + 1)
        lays = {q: Q.initialize(s, fl, fr, q, False, True) for q in labels}
        vecs = {q: torch.zeros(max(lays[q].total, 1), dtype=torch.float64,
                               device=self.dev) for q in labels}
        ket = _info_set([lay])
        li, ri = _info_set(lblk.infos), _info_set(rblk.infos)
        lo, ro = _op_set(lblk), _op_set(rblk)
        for (c1, opdq), terms in groups.items():
            for q in vdqs[(c1, opdq)]:
                if lays[q].total == 0:
                    continue
                pb = Problem(s, opdq, 0, ket, _info_set([lays[q]]), li, ri,
                             lo, ro, [t[0] for t in terms],
                             [t[1] for t in terms], [t[2] for t in terms],
                             [t[3] for t in terms], dict(site=i))
                h = EffectiveHamiltonian(pb, self.ctx, fp32=False,
                                         device_data=(lblk.data, rblk.data))
                try:
                    h(x, vecs[q][:lays[q].total])
                finally:
                    h.deallocate()
        # scale_perturbative_noise: each normalised, then all together to a
        # squared norm of `noise`
        for q in labels:
            nrm = float(torch.linalg.norm(vecs[q]))
            if nrm > 1e-20:
                vecs[q] /= nrm
        tot = float(torch.sqrt(sum(torch.dot(vecs[q], vecs[q])
                                   for q in labels))) if labels else 0.0
        if tot > 1e-20:
            for q in labels:
                vecs[q] *= np.sqrt(noise) / tot
        out = []
        for q in labels[1:]:             # density_matrix adds pkets 1..n-1
            ly = lays[q]
            out.append((T.BlockInfo(ly.quanta, ly.nbra, ly.nket, ly.shift,
                                    ly.dq, 0), vecs[q][:ly.total]))
        return out

    def _relieve_memory(self):
        """Under device-memory pressure (less than 45 % of the device free:
        one step of an n=40, M=3000 ab initio sweep allocates about 100 GB)
        drops the plans kept across sweeps (rebuilt on demand) and hands
        torch's cached blocks back to the device, so that library
        allocations can use them."""
        free, total = torch.cuda.mem_get_info(self.dev)
        if free >= 0.45 * total:
            return
        self.close()
        torch.cuda.empty_cache()

    def update_two_dot(self, i, forward, bond_dim, noise, conv):
        s, n = self.sym, self.n
        self._relieve_memory()
        t0 = tp = time.perf_counter()
        # the two-site wavefunction at (i, i+1)
        if self.tensors[i + 1] is not None:
            lay = self._wfn_layout(i)
            a, b = self.tensors[i], self.tensors[i + 1]
            if a.lay.wfn:   # wfn x rot
                psi = self._contract_wfn(a, True, b, lay)
            else:           # rot x wfn
                psi = self._contract_wfn(a, False, b, lay)
            self.tensors[i], self.tensors[i + 1] = psi, None
        psi = self.tensors[i]
        lay = psi.lay
        tp = self._tick("wavefunction", tp)
        lblk = self.left_block(i, self.lenv[i])
        rblk = self.right_block(i + 1, self.renv[i + 2])
        tp = self._tick("blocking", tp)
        h = self._eff_ham(i, lblk, rblk, lay)
        tp = self._tick("setup", tp)
        x = psi.data[:lay.total].clone()
        e, ndav, nflop, tdav = h.eigs(
            x, conv_thrd=conv, max_iter=self.davidson_max_iter,
            soft_max_iter=self.davidson_soft_max_iter,
            deflation_min_size=self.davidson_def_min,
            deflation_max_size=self.davidson_def_max,
            precision=self.precision, refine=True)
        h.deallocate()
        tp = self._tick("davidson", tp)
        self.davidson_macs += nflop
        energy = e + self.const_e
        bi = T.BlockInfo(lay.quanta, lay.nbra, lay.nket, lay.shift, lay.dq,
                         int(lay.fermion))
        if self.perturbative and noise:
            extra = self._perturbative_noise(i, lblk, rblk, lay, x, forward,
                                             noise)
            tp = self._tick("perturbative", tp)
            sp = T.split_density_matrix(
                self.ctx, s, bi, x, int(bond_dim), self.cutoff, forward,
                True, T.TRUNC_PHYSICAL, 1.0, extra=extra)
        else:
            sp = T.split_density_matrix(
                self.ctx, s, bi, x, int(bond_dim), self.cutoff, forward,
                True, T.TRUNC_PHYSICAL, 1.0, None,
                noise if noise and not self.perturbative else 0.0,
                self.noise_type, 0, rng=self.rng if noise else None)
        tp = self._tick("truncation", tp)
        (linfo, ldata), (rinfo, rdata) = sp.left, sp.right
        if forward:
            left = MPSTensor(_layout_of_split(linfo, False), ldata)
            right = MPSTensor(_layout_of_split(rinfo, True), rdata)
            self.left_dims[i + 1] = Q.extract_state_info(s, left.lay, True)
            env = self.rotate(lblk, left, self.left_dims[i + 1], False)
            if self._schemer_left(i):
                sc = self.schemer
                env = self.transform(env, sc["left_new_operator_names"],
                                     sc["left_new_operator_exprs"])
            self.lenv[i + 1] = env
            tp = self._tick("renormalisation", tp)
            self.tensors[i], self.tensors[i + 1] = left, right
            if i + 1 != n - 1:
                self.tensors[i + 1] = self._swap(right, i + 1, True)
        else:
            left = MPSTensor(_layout_of_split(linfo, True), ldata)
            right = MPSTensor(_layout_of_split(rinfo, False), rdata)
            self.right_dims[i + 1] = Q.extract_state_info(s, right.lay, False)
            env = self.rotate(rblk, right, self.right_dims[i + 1], True)
            if self._schemer_right(i + 1):
                sc = self.schemer
                env = self.transform(env, sc["right_new_operator_names"],
                                     sc["right_new_operator_exprs"])
            self.renv[i + 1] = env
            tp = self._tick("renormalisation", tp)
            self.tensors[i], self.tensors[i + 1] = left, right
            if i != 0:
                self.tensors[i] = self._swap(left, i, False)
        self._tick("wavefunction", tp)
        # the environment on the far side of the step is rebuilt by the next
        # sweep before it is read again: released now, so that one set of
        # environments is held on the device rather than two
        if forward:
            self.renv[i + 2] = None
        else:
            self.lenv[i] = None
        self.bonds.append(self.left_dims[i + 1] if forward else
                          self.right_dims[i + 1])
        return energy, sp.error, ndav, time.perf_counter() - t0

    def sweep(self, forward, bond_dim, noise, conv):
        n = self.n
        rng = range(self.center, n - 1) if forward else \
            range(self.center, -1, -1)
        es, errs = [], []
        for i in rng:
            e, err, nd, dt = self.update_two_dot(i, forward, bond_dim, noise,
                                                 conv)
            es.append(e)
            errs.append(err)
            self.steps.append((self.isweep, len(es) - 1, e, err, nd, dt))
            self.center = i
        if forward:
            self.center = n - 2
        else:
            self.center = 0
        return min(es), max(errs)

    def solve(self, n_sweeps, forward=True, tol=1e-6, davidson_conv=None):
        """DMRG::solve: bond dimensions, noises and Davidson thresholds
        extended to n_sweeps as block2 extends them; stops when the energy
        change is below tol with the final noise and bond dimension."""
        bd = self.bond_dims + [self.bond_dims[-1]] * max(
            0, n_sweeps - len(self.bond_dims))
        ns = self.noises + [self.noises[-1] if self.noises else 0.0] * max(
            0, n_sweeps - len(self.noises))
        dc = list(davidson_conv or [])
        for i in range(len(dc), len(ns)):
            dc.append((ns[i] if ns[i] != 0 else (tol if tol != 0 else 1e-9))
                      * 0.1)
        self.davidson_conv_thrds = dc
        self.energies, self.discarded_weights, self.sweep_time = [], [], []
        self.steps, self.bonds = [], []
        # seconds per phase of the sweep steps (host clock; the phases end
        # in synchronisations), and the Davidson's multiply-adds
        self.phase_time, self.davidson_macs = {}, 0
        self.init_environments()
        for iw in range(n_sweeps):
            self.isweep = iw
            t0 = time.perf_counter()
            e, dw = self.sweep(forward, bd[iw], ns[iw], dc[iw])
            torch.cuda.synchronize(self.dev)
            self.sweep_time.append(time.perf_counter() - t0)
            self.energies.append(e)
            self.discarded_weights.append(dw)
            conv = (len(self.energies) >= 2 and tol > 0 and
                    abs(self.energies[-1] - self.energies[-2]) < tol and
                    ns[iw] == ns[-1] and bd[iw] == bd[-1])
            forward = not forward
            if conv:
                break
        return self.energies[-1]
