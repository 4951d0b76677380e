"""Quantum labels, StateInfo fusion and block layouts on the host.

Labels are block2's packed 64-bit words (SZLongLong / SU2LongLong,
src/core/symmetry.hpp:656-732, 1184-1290):
  SZ : bits 48..63 N, 16..31 2Sz (signed), 0..15 point group
  SU2: bits 48..63 N, 32..47 2S of the bra (composite labels) or the low
       end of a spin range, 16..31 2S, 0..15 point group
The functions follow the label algebra the library's plan builder uses
(csrc/symmetry.hpp); StateInfo and SparseMatrixInfo construction follow
  StateInfo::tensor_product(a, b, cref)   src/core/state_info.hpp:231-254
  SparseMatrixInfo::initialize            src/core/sparse_matrix.hpp:727-757
  SparseMatrixInfo::extract_state_info    src/core/sparse_matrix.hpp:759-781
"""
import numpy as np

SZ, SU2 = 0, 1
INVALID = (1 << 64) - 1
_M16 = 0xFFFF


def q_n(q):
    v = (q >> 48) & _M16
    return v - 0x10000 if v & 0x8000 else v


def q_pg(q):
    return q & _M16


def q_twos(s, q):
    v = (q >> 16) & _M16
    if s == SZ and v & 0x8000:
        v -= 0x10000
    return v


def q_twos_low(q):
    return (q >> 32) & _M16


def q_make(s, n, twos_low, twos, pg):
    w = ((n & _M16) << 48) | ((twos & _M16) << 16) | (pg & _M16)
    if s == SU2:
        w |= (twos_low & _M16) << 32
    return w


def q_multiplicity(s, q):
    return 1 if s == SZ else q_twos(s, q) + 1


def q_is_fermion(q):
    return bool((q >> 16) & 1)


def q_neg(s, q):
    if s == SZ:
        return q_make(s, -q_n(q), 0, -q_twos(s, q), q_pg(q))
    return q_make(s, -q_n(q), q_twos_low(q), q_twos(s, q), q_pg(q))


def q_add(s, a, b):
    """a + b; for SU2 the spin range |Sa - Sb| .. Sa + Sb (ranges added
    as block2 adds them)."""
    n, pg = q_n(a) + q_n(b), q_pg(a) ^ q_pg(b)
    if s == SZ:
        return q_make(s, n, 0, q_twos(s, a) + q_twos(s, b), pg)
    ta, tb = q_twos(s, a), q_twos(s, b)
    la, lb = q_twos_low(a), q_twos_low(b)
    # the lowest spin: the smaller non-negative of 2Sa - 2Sb_low and
    # 2Sb - 2Sa_low, compared in the 16-bit field with wrap-around
    lr, rl = (ta - lb) & _M16, (tb - la) & _M16
    return q_make(s, n, min(lr, rl), ta + tb, pg)


def q_count(s, q):
    return 1 if s == SZ else ((q_twos(s, q) - q_twos_low(q)) >> 1) + 1


def q_at(s, q, i):
    if s == SZ:
        return q
    t = q_twos_low(q) + 2 * i
    return q_make(s, q_n(q), t, t, q_pg(q))


def q_members(s, q):
    return [q_at(s, q, i) for i in range(q_count(s, q))]


def q_get_ket(s, q):
    if s == SZ:
        return q
    t = q_twos(s, q)
    return q_make(s, q_n(q), t, t, q_pg(q))


def q_get_bra(s, q, dq):
    if s == SZ:
        return q_add(s, q, dq)
    t = q_twos_low(q)
    return q_make(s, q_n(q) + q_n(dq), t, t, q_pg(q) ^ q_pg(dq))


def _triangle(a, b, c):
    return not ((a + b + c) & 1) and abs(a - b) <= c <= a + b


def q_combine(s, dq, bra, ket):
    """Composite label of the block <bra| op(dq) |ket>, INVALID if the
    symmetry forbids it."""
    if s == SZ:
        return ket if q_add(s, ket, dq) == bra else INVALID
    k = q_make(s, q_n(ket), q_twos(s, bra), q_twos(s, ket), q_pg(ket))
    if q_get_bra(s, k, dq) != bra or not _triangle(
            q_twos(s, k), q_twos(s, dq), q_twos(s, bra)):
        return INVALID
    return k


class StateInfo:
    """Quanta (ascending uint64) with their state counts."""

    def __init__(self, quanta, n_states):
        self.q = np.asarray(quanta, dtype=np.uint64)
        self.n = np.asarray(n_states, dtype=np.uint32)
        assert self.q.shape == self.n.shape

    def __len__(self):
        return len(self.q)

    @property
    def total(self):
        return int(self.n.sum())

    def find(self, q):
        i = int(np.searchsorted(self.q, np.uint64(q)))
        return i if i < len(self.q) and int(self.q[i]) == q else -1

    @staticmethod
    def from_dict(d):
        return StateInfo(d["quanta"], d["n_states"])

    def __eq__(self, o):
        return (isinstance(o, StateInfo) and np.array_equal(self.q, o.q)
                and np.array_equal(self.n, o.n))

    def __repr__(self):
        return "StateInfo(%d quanta, %d states)" % (len(self), self.total)


def tensor_product(s, a, b, cref):
    """StateInfo::tensor_product(a, b, cref): the quanta of cref, each with
    the summed dimensions of the (a, b) pairs coupling to it; empty ones
    removed."""
    cap = 0xFFFFFFFF
    n = np.zeros(len(cref), dtype=np.uint64)
    for i in range(len(a)):
        qa, na = int(a.q[i]), int(a.n[i])
        for j in range(len(b)):
            qc = q_add(s, qa, int(b.q[j]))
            for k in range(q_count(s, qc)):
                ic = cref.find(q_at(s, qc, k))
                if ic != -1:
                    n[ic] = min(int(n[ic]) + na * int(b.n[j]), cap)
    keep = n != 0
    return StateInfo(cref.q[keep], n[keep].astype(np.uint32))


class Layout:
    """A SparseMatrixInfo: blocks (sorted composite quanta) with their
    row / column dimensions and element offsets."""

    def __init__(self, quanta, nbra, nket, dq=0, fermion=False, wfn=False):
        self.quanta = np.asarray(quanta, dtype=np.uint64)
        self.nbra = np.asarray(nbra, dtype=np.uint32)
        self.nket = np.asarray(nket, dtype=np.uint32)
        self.dq = int(dq)
        self.fermion = bool(fermion)
        self.wfn = bool(wfn)
        sz = self.nbra.astype(np.uint64) * self.nket.astype(np.uint64)
        self.shift = np.concatenate([[0], np.cumsum(sz)[:-1]]).astype(
            np.uint64) if len(sz) else np.zeros(0, np.uint64)
        self.total = int(sz.sum()) if len(sz) else 0

    def __len__(self):
        return len(self.quanta)

    def find(self, q):
        i = int(np.searchsorted(self.quanta, np.uint64(q)))
        return i if i < len(self.quanta) and int(self.quanta[i]) == q else -1

    @staticmethod
    def from_dict(d, wfn=None):
        """From a parsed SparseMatrixInfo (effprob._Reader.info)."""
        lay = Layout(d["quanta"], d["nbra"], d["nket"], d["dq"],
                     d["fermion"], bool(d["flags"] & 1) if wfn is None
                     else wfn)
        if not np.array_equal(lay.shift, np.asarray(d["shift"], np.uint64)):
            raise ValueError("block offsets differ from the packed layout")
        return lay

    def same(self, o):
        return (np.array_equal(self.quanta, o.quanta)
                and np.array_equal(self.nbra, o.nbra)
                and np.array_equal(self.nket, o.nket) and self.dq == o.dq)


def initialize(s, bra, ket, dq, fermion, wfn=False):
    """SparseMatrixInfo::initialize(bra, ket, dq, is_fermion, wfn)."""
    qs = []
    for i in range(len(ket)):
        q = q_neg(s, int(ket.q[i])) if wfn else int(ket.q[i])
        bs = q_add(s, dq, q)
        for k in range(q_count(s, bs)):
            b = q_at(s, bs, k)
            if bra.find(b) != -1:
                qs.append(q_combine(s, dq, b, q))
    qs.sort()
    nb, nk = [], []
    for x in qs:
        kq = q_get_ket(s, x)
        ik = ket.find(q_neg(s, kq) if wfn else kq)
        ib = bra.find(q_get_bra(s, x, dq))
        assert ik != -1 and ib != -1
        nk.append(int(ket.n[ik]))
        nb.append(int(bra.n[ib]))
    return Layout(qs, nb, nk, dq, fermion, wfn)


def extract_state_info(s, lay, right):
    """SparseMatrixInfo::extract_state_info: the row (right False) or
    column basis of a block layout."""
    vac = lay.dq == q_at(s, q_add(s, lay.dq, q_neg(s, lay.dq)), 0)
    if vac and not lay.wfn:
        return StateInfo(lay.quanta.copy(),
                         (lay.nket if right else lay.nbra).copy())
    qs = {}
    for i in range(len(lay)):
        x = int(lay.quanta[i])
        if right:
            k = q_get_ket(s, x)
            key = q_neg(s, k) if lay.wfn else k
            qs[key] = int(lay.nket[i])
        else:
            qs[q_get_bra(s, x, lay.dq)] = int(lay.nbra[i])
    keys = sorted(qs)
    return StateInfo(keys, [qs[k] for k in keys])
