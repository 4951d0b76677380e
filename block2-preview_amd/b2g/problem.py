"""Block-sparse effective-Hamiltonian problem (host side, numpy).

Holds what block2's EffectiveHamiltonian constructor receives
(src/dmrg/effective_hamiltonian.hpp:126-224): the ket/bra SparseMatrixInfo,
left/right operator infos keyed by delta quantum, the operator blocks behind
lopt/ropt->ops and the strings of the H_eff OpSum.  Arrays are flattened
structure-of-arrays:

  infos: dq[i], fermion[i], n[i], off[i]; block j of info i is entry off[i]+j
         of quanta / nbra / nket / shift (element offset inside the operator)
  ops:   dq[k], factor[k], info[k], off[k] (element offset in the data buffer)
  terms: conj[t], a[t], b[t], factor[t]   (H += factor op(L[a]) x op(R[b]))
"""
import ctypes as C
import json

import numpy as np

from . import _lib as L


class InfoSet:
    """A list of block-sparse infos in flattened form."""

    def __init__(self, dq, fermion, n, off, quanta, nbra, nket, shift):
        self.dq = np.ascontiguousarray(dq, dtype=np.uint64)
        self.fermion = np.ascontiguousarray(fermion, dtype=np.uint32)
        self.n = np.ascontiguousarray(n, dtype=np.uint32)
        self.off = np.ascontiguousarray(off, dtype=np.uint32)
        self.quanta = np.ascontiguousarray(quanta, dtype=np.uint64)
        self.nbra = np.ascontiguousarray(nbra, dtype=np.uint32)
        self.nket = np.ascontiguousarray(nket, dtype=np.uint32)
        self.shift = np.ascontiguousarray(shift, dtype=np.uint64)

    def __len__(self):
        return len(self.dq)

    def total(self, i):
        n = int(self.n[i])
        if n == 0:
            return 0
        o = int(self.off[i]) + n - 1
        return int(self.shift[o]) + int(self.nbra[o]) * int(self.nket[o])

    def block(self, i, j):
        o = int(self.off[i]) + j
        return int(self.shift[o]), int(self.nbra[o]), int(self.nket[o])


class OpSet:
    def __init__(self, dq, factor, info, off, data, flags=None):
        self.dq = np.ascontiguousarray(dq, dtype=np.uint64)
        self.factor = np.ascontiguousarray(factor, dtype=np.float64)
        self.info = np.ascontiguousarray(info, dtype=np.int32)
        self.off = np.ascontiguousarray(off, dtype=np.uint64)
        # bit 0: structure only (a delayed operator never formed)
        self.flags = (np.zeros(len(self.dq), np.int32) if flags is None else
                      np.ascontiguousarray(flags, dtype=np.int32))
        # None: the operator data live only on the device (see to_c)
        self.data = (None if data is None else
                     np.ascontiguousarray(data, dtype=np.float64))

    def __len__(self):
        return len(self.dq)


class Problem:
    """One two-site effective Hamiltonian (see module docstring)."""

    def __init__(self, sym, opdq, hop_left_vacuum, ket, bra, linfo, rinfo,
                 lops, rops, term_conj, term_a, term_b, term_factor,
                 meta=None):
        self.sym = int(sym)  # 0 SZ, 1 SU2
        self.opdq = int(opdq)
        self.hop_left_vacuum = int(hop_left_vacuum)
        self.ket = ket  # InfoSet of length 1
        self.bra = bra
        self.linfo, self.rinfo = linfo, rinfo
        self.lops, self.rops = lops, rops
        self.term_conj = np.ascontiguousarray(term_conj, dtype=np.uint8)
        self.term_a = np.ascontiguousarray(term_a, dtype=np.int32)
        self.term_b = np.ascontiguousarray(term_b, dtype=np.int32)
        self.term_factor = np.ascontiguousarray(term_factor, dtype=np.float64)
        self.meta = meta or {}
        self._keep = []
        # delayed (three-tensor) strings: term_kind 1 uses sub-operators
        # dlops[term_da], drops[term_db] of the delayed tensor on side
        # delayed_side (1 left, 2 right; their data sit in that side's
        # buffer) and the tensor-product lists tp[i] of that side's infos
        nt = len(self.term_conj)
        self.term_kind = np.zeros(nt, np.uint8)
        self.term_da = np.full(nt, -1, np.int32)
        self.term_db = np.full(nt, -1, np.int32)
        self.term_dconj = np.zeros(nt, np.uint8)
        self.delayed_side = 0
        self.dlinfo = self.drinfo = self.dlops = self.drops = None
        self.tp = []

    @property
    def ket_len(self):
        return self.ket.total(0)

    @property
    def bra_len(self):
        return self.bra.total(0)

    # ------------------------------------------------------------ fixtures
    @staticmethod
    def from_npz(path):
        """Load a problem dumped by the reference harness
        (tests/golden/*.npz)."""
        z = np.load(path, allow_pickle=False)
        meta = json.loads(bytes(z["meta_json"]).decode())

        def info(p):
            return InfoSet(z[p + "_dq"], z[p + "_fermion"], z[p + "_n"],
                           z[p + "_off"], z[p + "_quanta"], z[p + "_nbra"],
                           z[p + "_nket"], z[p + "_shift"].astype(np.uint64))

        def ops(p):
            return OpSet(z[p + "op_dq"], z[p + "op_factor"], z[p + "op_info"],
                         z[p + "op_off"], z[p + "op_data"])

        sym = 1 if meta["sym"] == "su2" else 0
        lops, rops = ops("l"), ops("r")
        for o, p in ((lops, "l"), (rops, "r")):
            if p + "op_flags" in z.files:
                o.flags = np.ascontiguousarray(z[p + "op_flags"], np.int32)
        side = int(meta.get("delayed_side", 0))
        if side:
            # sub-operator data appended to the delayed side's buffer
            host = lops if side == 1 else rops
            base = len(host.data)
            dl, dr = ops("dl"), ops("dr")
            dl.off = dl.off + np.uint64(base)
            dr.off = dr.off + np.uint64(base + len(dl.data))
            host.data = np.concatenate([host.data, dl.data, dr.data])
        pb = Problem(sym, int(meta["opdq"]), int(meta["hop_left_vacuum"]),
                     info("ket"), info("bra"), info("linfo"), info("rinfo"),
                     lops, rops, z["term_conj"], z["term_a"],
                     z["term_b"], z["term_factor"], meta)
        if "term_kind" in z.files:
            pb.term_kind = np.ascontiguousarray(z["term_kind"], np.uint8)
            pb.term_da = np.ascontiguousarray(z["term_da"], np.int32)
            pb.term_db = np.ascontiguousarray(z["term_db"], np.int32)
            pb.term_dconj = np.ascontiguousarray(z["term_dconj"], np.uint8)
        if side:
            pb.delayed_side = side
            pb.dlinfo, pb.drinfo = info("dlinfo"), info("drinfo")
            pb.dlops, pb.drops = dl, dr
            tn = z["tp_n"].reshape(-1, 5)
            tnc = z["tp_nc"]
            qo = co = 0
            for i in range(len(tnc)):
                nq = max(int(tn[i, 4]), 0)
                nc = int(tnc[i])
                pb.tp.append(dict(
                    n=tn[i].astype(np.int32), nc=nc,
                    quanta=np.ascontiguousarray(z["tp_quanta"][qo:qo + nq]),
                    idx=np.ascontiguousarray(z["tp_idx"][qo:qo + nq]),
                    ia=np.ascontiguousarray(z["tp_ia"][co:co + nc]),
                    ib=np.ascontiguousarray(z["tp_ib"][co:co + nc]),
                    ic=np.ascontiguousarray(z["tp_ic"][co:co + nc]),
                    stride=np.ascontiguousarray(z["tp_stride"][co:co + nc]),
                    factor=np.ascontiguousarray(z["tp_factor"][co:co + nc])))
                qo += nq
                co += nc
        pb.golden = {k: z[k] for k in z.files}
        return pb

    # --------------------------------------------------------- C structure
    def _sparse_info(self, s, i):
        o, n = int(s.off[i]), int(s.n[i])
        q = np.ascontiguousarray(s.quanta[o:o + n])
        nb = np.ascontiguousarray(s.nbra[o:o + n])
        nk = np.ascontiguousarray(s.nket[o:o + n])
        sh = np.ascontiguousarray(s.shift[o:o + n])
        self._keep += [q, nb, nk, sh]
        return L.b2g_sparse_info(int(s.dq[i]), int(s.fermion[i]), n,
                                 L.ptr(q, C.c_uint64), L.ptr(nb, C.c_uint32),
                                 L.ptr(nk, C.c_uint32),
                                 L.ptr(sh, C.c_uint64))

    def to_c(self, device_data=None):
        """b2g_problem structure pointing into this object's arrays.
        device_data=(left_ptr, left_len, right_ptr, right_len): operator
        buffers already on the device (pass B2G_EFFHAM_DEVICE_DATA)."""
        self._keep = []
        lin = (L.b2g_sparse_info * max(len(self.linfo), 1))(
            *[self._sparse_info(self.linfo, i) for i in range(len(self.linfo))])
        rin = (L.b2g_sparse_info * max(len(self.rinfo), 1))(
            *[self._sparse_info(self.rinfo, i) for i in range(len(self.rinfo))])
        def op_arr(o):
            return (L.b2g_operator * max(len(o), 1))(
                *[L.b2g_operator(int(o.dq[k]), float(o.factor[k]),
                                 int(o.info[k]), int(o.flags[k]),
                                 int(o.off[k])) for k in range(len(o))])
        lo, ro = op_arr(self.lops), op_arr(self.rops)
        nt = len(self.term_conj)
        tm = (L.b2g_term * max(nt, 1))(
            *[L.b2g_term(a=int(self.term_a[t]), b=int(self.term_b[t]),
                         conj=int(self.term_conj[t]),
                         kind=int(self.term_kind[t]),
                         factor=float(self.term_factor[t]),
                         da=int(self.term_da[t]), db=int(self.term_db[t]),
                         dconj=int(self.term_dconj[t]))
              for t in range(nt)])
        self._keep += [lin, rin, lo, ro, tm]
        p = L.b2g_problem()
        p.symmetry = self.sym
        p.n_left_infos, p.n_right_infos = len(self.linfo), len(self.rinfo)
        p.n_left_ops, p.n_right_ops, p.n_terms = len(self.lops), len(self.rops), nt
        p.opdq, p.hop_left_vacuum = self.opdq, self.hop_left_vacuum
        p.ket = self._sparse_info(self.ket, 0)
        p.bra = self._sparse_info(self.bra, 0)
        p.left_info_dq = L.ptr(self.linfo.dq, C.c_uint64)
        p.right_info_dq = L.ptr(self.rinfo.dq, C.c_uint64)
        p.left_infos, p.right_infos = lin, rin
        p.left_ops, p.right_ops, p.terms = lo, ro, tm
        if device_data is None and self.lops.data is None:
            # structure only (host-only plans): no data, declared lengths
            dl = getattr(self, "data_len", {0: 0, 1: 0})
            p.left_data = p.right_data = None
            p.left_len, p.right_len = dl[0], dl[1]
        elif device_data is None:
            p.left_data = L.ptr(self.lops.data, C.c_double)
            p.right_data = L.ptr(self.rops.data, C.c_double)
            p.left_len, p.right_len = len(self.lops.data), len(self.rops.data)
        else:
            lp, ll, rp, rl = device_data
            p.left_data = C.cast(C.c_void_p(lp), L.f64p)
            p.right_data = C.cast(C.c_void_p(rp), L.f64p)
            p.left_len, p.right_len = ll, rl
        p.delayed_side = self.delayed_side
        if self.delayed_side:
            dli = (L.b2g_sparse_info * max(len(self.dlinfo), 1))(
                *[self._sparse_info(self.dlinfo, i)
                  for i in range(len(self.dlinfo))])
            dri = (L.b2g_sparse_info * max(len(self.drinfo), 1))(
                *[self._sparse_info(self.drinfo, i)
                  for i in range(len(self.drinfo))])
            dlo, dro = op_arr(self.dlops), op_arr(self.drops)
            tps = (L.b2g_tp_conn * max(len(self.tp), 1))()
            for i, t in enumerate(self.tp):
                c = tps[i]
                for k in range(5):
                    c.n[k] = int(t["n"][k])
                c.nc = int(t["nc"])
                if c.nc > 0:
                    arrs = [np.ascontiguousarray(t[k], dt) for k, dt in (
                        ("quanta", np.uint64), ("idx", np.uint32),
                        ("ia", np.uint32), ("ib", np.uint32),
                        ("ic", np.uint32), ("stride", np.uint64),
                        ("factor", np.float64))]
                    self._keep += arrs
                    c.quanta = L.ptr(arrs[0], C.c_uint64)
                    c.idx = L.ptr(arrs[1], C.c_uint32)
                    c.ia = L.ptr(arrs[2], C.c_uint32)
                    c.ib = L.ptr(arrs[3], C.c_uint32)
                    c.ic = L.ptr(arrs[4], C.c_uint32)
                    c.stride = L.ptr(arrs[5], C.c_uint64)
                    c.factor = L.ptr(arrs[6], C.c_double)
            self._keep += [dli, dri, dlo, dro, tps]
            p.n_dl_infos, p.n_dr_infos = len(self.dlinfo), len(self.drinfo)
            p.n_dl_ops, p.n_dr_ops = len(self.dlops), len(self.drops)
            p.dl_infos, p.dr_infos = dli, dri
            p.dl_ops, p.dr_ops = dlo, dro
            p.tp = tps
        self._keep.append(p)
        return p
