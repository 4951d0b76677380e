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
    def __init__(self, dq, factor, info, off, data):
        self.dq = np.ascontiguousarray(dq, dtype=np.uint64)
        self.factor = np.ascontiguousarray(factor, dtype=np.float64)
        self.info = np.ascontiguousarray(info, dtype=np.int32)
        self.off = np.ascontiguousarray(off, dtype=np.uint64)
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
        pb = Problem(sym, int(meta["opdq"]), int(meta["hop_left_vacuum"]),
                     info("ket"), info("bra"), info("linfo"), info("rinfo"),
                     ops("l"), ops("r"), z["term_conj"], z["term_a"],
                     z["term_b"], z["term_factor"], meta)
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
        lo = (L.b2g_operator * max(len(self.lops), 1))(
            *[L.b2g_operator(int(self.lops.dq[k]), float(self.lops.factor[k]),
                             int(self.lops.info[k]), 0, int(self.lops.off[k]))
              for k in range(len(self.lops))])
        ro = (L.b2g_operator * max(len(self.rops), 1))(
            *[L.b2g_operator(int(self.rops.dq[k]), float(self.rops.factor[k]),
                             int(self.rops.info[k]), 0, int(self.rops.off[k]))
              for k in range(len(self.rops))])
        nt = len(self.term_conj)
        tm = (L.b2g_term * max(nt, 1))(
            *[L.b2g_term(int(self.term_a[t]), int(self.term_b[t]),
                         int(self.term_conj[t]), 0, float(self.term_factor[t]))
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
        self._keep.append(p)
        return p
