"""block2's EFFPROB1 effective-problem files, read into a Problem.

EffectiveHamiltonian::save_eff_problem (src/dmrg/effective_hamiltonian.hpp:
254-261) writes, through EffectiveProblem::save_problem_file
(src/core/effective_problem.hpp:72-243), everything that defines one
two-site H*psi except the operator numbers: the DelayedOperatorTensor with
its H expression and the left/right OperatorTensors (every operator's
SparseMatrixInfo and ConnectionInfo, its element count), the ket and bra
infos, the operator quantum and the wavefunction ConnectionInfos.  This
module parses such a file and builds the Problem the device engine takes,
structure only: the operator data (which the file does not hold) are
attached afterwards, in the order Problem.lops / rops lists the operators,
by the caller or from block2's OperatorTensor files of the two
environments (load_operator_tensor, attach_operator_data).

Serialisations followed:
  EffectiveProblem header / tensors / metadata  effective_problem.hpp:72-243
  OperatorTensor ops and lmat/rmat markers      operator_tensor.hpp:78-140
  OpExpr::save (Zero, Elem, ElemRef, Prod,      src/core/expr.hpp:209-216,
    SumProd, Sum, ExprRef)                        441-499, 584-612, 744-800,
                                                  869-885, 931-950
  save_symbolic                                 src/core/symbolic.hpp:164-215
  SparseMatrixInfo::save_data                   src/core/sparse_matrix.hpp:555-575
  SparseMatrix::save_data (operator numbers)    src/core/sparse_matrix.hpp:954-973
  DelayedSparseMatrix::save_data                src/core/delayed_sparse_matrix.hpp:94-103

Operators are numbered the way the effective Hamiltonian meets them: by
first appearance in the H expression's strings (left operator a, right
operator b; the sub-operators of SumProd strings in a delayed tensor on
their own lists).  Operator factors are not part of the file (block2 keeps
operators normalised, factor 1) and are read as 1.  An operator of a
delayed tensor that a plain product string uses and that has an expression
(DelayedOperatorTensor::exprs) is formed before the multiply, so it gets
data; so does an operator of a delayed tensor that was allocated (a
nonzero element count in the file); the other operators of a delayed
tensor are structure only.
"""
import struct

import numpy as np

from .problem import InfoSet, OpSet, Problem

MAGIC = b"EFFPROB1"
_SYM = {1: 1, 2: 0}  # file tag (SU2 = 1, SZ = 2) -> Problem.sym
ZERO, ELEM, PROD, SUM, ELEMREF, SUMPROD, EXPRREF = range(7)
NORMAL, CSR, ARCHIVED, DELAYED = range(4)


class FormatError(ValueError):
    pass


class _Reader:
    def __init__(self, buf, ubond_size=4):
        self.b = memoryview(buf)
        self.p = 0
        self.ubond = ubond_size

    def take(self, n):
        if n < 0 or self.p + n > len(self.b):
            raise FormatError("file ends inside an EFFPROB1 record")
        v = self.b[self.p:self.p + n]
        self.p += n
        return v

    def u8(self):
        return self.take(1)[0]

    def i8(self):
        return struct.unpack("<b", self.take(1))[0]

    def i32(self):
        return struct.unpack("<i", self.take(4))[0]

    def u16(self):
        return struct.unpack("<H", self.take(2))[0]

    def u32(self):
        return struct.unpack("<I", self.take(4))[0]

    def u64(self):
        return struct.unpack("<Q", self.take(8))[0]

    def f64(self):
        return struct.unpack("<d", self.take(8))[0]

    def arr(self, dtype, n):
        dt = np.dtype(dtype)
        return np.frombuffer(self.take(dt.itemsize * n), dtype=dt).copy()

    # ---------------------------------------------------------- expressions
    def expr(self):
        tp = self.u8()
        if tp == ZERO:
            return ("zero",)
        if tp == ELEM:
            name, site = self.u8(), self.u64()
            factor, q = self.f64(), self.u64()
            return ("elem", name, site, factor, q)
        if tp == ELEMREF:
            factor, trans = self.i8(), self.i8()
            return ("elemref", factor, trans, self.expr())
        if tp == PROD:
            factor, conj, has = self.f64(), self.u8(), self.u8()
            a = self.expr() if has & 1 else None
            b = self.expr() if has & 2 else None
            return ("prod", factor, conj, a, b)
        if tp == SUMPROD:
            factor, conj, has = self.f64(), self.u8(), self.u8()
            a = self.expr() if has & 1 else None
            b = self.expr() if has & 2 else None
            c = self.expr() if has & 4 else None
            n = self.i32()
            ops = [self.expr() for _ in range(n)]
            conjs = [self.u8() for _ in range(n)]
            return ("sumprod", factor, conj, a, b, c, ops, conjs)
        if tp == SUM:
            n = self.i32()
            return ("sum", [self.expr() for _ in range(n)])
        if tp == EXPRREF:
            local = self.u8()
            op = self.expr()
            orig = self.expr() if self.u8() & 1 else None
            return ("exprref", local, op, orig)
        raise FormatError("unknown expression type %d" % tp)

    def symbolic(self):
        tp, m, n, sz = self.u8(), self.i32(), self.i32(), self.i32()
        data = [self.expr() for _ in range(sz)]
        idx = self.arr(np.int32, 2 * sz).reshape(sz, 2) if tp == 2 else None
        return dict(type=tp, m=m, n=n, data=data, indices=idx)

    # ---------------------------------------------------- sparse matrices
    def info(self):
        dq, n = self.u64(), self.i32()
        if n < 0 or n >= (1 << 28):
            raise FormatError("bad block count %d in sparse-matrix info" % n)
        quanta = self.arr(np.uint64, n)
        dw = {4: 2 * n, 2: n, 1: n - (n >> 1)}[self.ubond]
        dims = self.take(4 * dw)
        ut = {4: np.uint32, 2: np.uint16, 1: np.uint8}[self.ubond]
        d = np.frombuffer(dims, dtype=ut)[:2 * n].astype(np.uint32)
        shift = self.arr(np.uint32, n).astype(np.uint64)
        fermion, flags = self.u8(), self.u8()
        return dict(dq=dq, n=n, quanta=quanta, nbra=d[:n].copy(),
                    nket=d[n:2 * n].copy(), shift=shift, fermion=fermion,
                    flags=flags)

    def cinfo(self):
        n = self.arr(np.int32, 5)
        nc = self.i32()
        c = dict(n=n, nc=nc)
        nq = int(n[4])
        if nq != 0:
            c["quanta"] = self.arr(np.uint64, nq)
            c["idx"] = self.arr(np.uint32, nq)
        else:
            c["quanta"] = np.zeros(0, np.uint64)
            c["idx"] = np.zeros(0, np.uint32)
        if nc != 0:
            c["stride"] = self.arr(np.uint64, nc)
            c["factor"] = self.arr(np.float64, nc)
            c["ia"] = self.arr(np.uint32, nc)
            c["ib"] = self.arr(np.uint32, nc)
            c["ic"] = self.arr(np.uint32, nc)
        else:
            for k, dt in (("stride", np.uint64), ("factor", np.float64),
                          ("ia", np.uint32), ("ib", np.uint32),
                          ("ic", np.uint32)):
                c[k] = np.zeros(0, dt)
        return c

    def op_matrix(self):
        tp = self.u8()
        info = self.info()
        cinfo = self.cinfo() if self.u8() else None
        m = dict(type=tp, info=info, cinfo=cinfo, total=None, expr=None)
        if tp == NORMAL:
            m["total"] = self.u64()
        elif tp == DELAYED:
            m["m"] = self.u16()
            m["expr"] = self.expr()
        else:
            raise FormatError("operator matrix type %d (CSR / archived) "
                              "is not supported" % tp)
        return m

    def tensor(self):
        if not self.u8():
            return None
        tag = self.u8()
        lr = self.u8()
        t = dict(delayed=tag == 1, lmat=None, rmat=None)
        if lr in (1, 2):
            t["lmat"] = self.symbolic()
            if lr == 1:
                t["rmat"] = t["lmat"]
        elif lr == 3:
            t["rmat"] = self.symbolic()
        elif lr == 0:
            t["lmat"] = self.symbolic()
            t["rmat"] = self.symbolic()
        elif lr != 4:
            raise FormatError("bad operator-tensor marker %d" % lr)
        n = self.i32()
        t["ops"] = [(self.expr(), self.op_matrix()) for _ in range(n)]
        if t["delayed"]:
            t["dops"] = [self.expr() for _ in range(self.i32())]
            t["mat"] = self.symbolic() if self.u8() else None
            t["stacked_mat"] = self.symbolic() if self.u8() else None
            t["lopt"] = self.tensor()
            t["ropt"] = self.tensor()
            t["exprs"] = [(self.expr(), self.u64(), self.f64())
                          for _ in range(self.i32())]
        return t

    def data_tensor(self):
        """OperatorTensor::save_data(ofs) (operator_tensor.hpp:122-148):
        markers, symbolics, then per operator its key, matrix type,
        SparseMatrixInfo and SparseMatrix::save_data payload (factor,
        element count, elements)."""
        lr = self.u8()
        t = dict(lmat=None, rmat=None)
        if lr in (1, 2):
            t["lmat"] = self.symbolic()
            if lr == 1:
                t["rmat"] = t["lmat"]
        elif lr == 3:
            t["rmat"] = self.symbolic()
        elif lr == 0:
            t["lmat"] = self.symbolic()
            t["rmat"] = self.symbolic()
        elif lr != 4:
            raise FormatError("bad operator-tensor marker %d" % lr)
        ops = []
        for _ in range(self.i32()):
            key = self.expr()
            tp = self.u8()
            info = self.info()
            m = dict(type=tp, info=info, factor=1.0, data=None)
            if tp == NORMAL:
                m["factor"] = self.f64()
                total = self.u64()
                if total == (1 << 64) - 1:
                    raise FormatError("compressed operator storage is not "
                                      "supported")
                if total and total != _total(info):
                    raise FormatError("operator element count %d differs "
                                      "from its info (%d)"
                                      % (total, _total(info)))
                m["data"] = self.arr(np.float64, total)
            elif tp == DELAYED:
                m["m"] = self.u16()
                m["expr"] = self.expr()
            else:
                raise FormatError("operator matrix type %d (CSR / archived) "
                                  "is not supported" % tp)
            ops.append((key, m))
        t["ops"] = ops
        return t

    def metadata(self):
        if not self.u8():
            return None
        tp = self.u8()
        info = self.info()
        return dict(type=tp, info=info, factor=self.f64(), total=self.u64())


class EffProblemFile:
    """Parsed EFFPROB1 file: header fields, the effective Hamiltonian's
    delayed operator tensor `op`, ket (`cmat`) and bra (`vmat`) metadata,
    the operator quantum `opdq` and (operator_quanta[i], wfn_cinfos[i])."""

    def __init__(self, buf):
        r = _Reader(buf)
        if bytes(r.take(8)) != MAGIC:
            raise FormatError("not an EFFPROB1 file")
        self.version = r.u32()
        if self.version != 2:
            raise FormatError("EFFPROB1 version %d (2 is read)" % self.version)
        tag = r.u8()
        q_size, fl_size, fp_size, cpx_sz, ub, self.align = (
            r.u8(), r.u8(), r.u8(), r.u8(), r.u8(), r.u8())
        if tag not in _SYM or q_size != 8:
            raise FormatError("symmetry tag %d / quantum size %d: SU2 or SZ "
                              "with 8-byte quanta are read" % (tag, q_size))
        if fl_size != 8 or fp_size != 8 or cpx_sz != 1:
            raise FormatError("real FP64 problems are read (element size %d)"
                              % fl_size)
        if ub not in (1, 2, 4):
            raise FormatError("bad ubond size %d" % ub)
        r.ubond = ub
        self.sym = _SYM[tag]
        self.ubond_size = ub
        self.isite = r.i32()
        self.forward = bool(r.u8())
        self.op = r.tensor()
        if self.op is None or not self.op["delayed"]:
            raise FormatError("the effective operator is not a delayed "
                              "operator tensor")
        self.cmat = r.metadata()
        self.vmat = r.metadata()
        self.opdq = r.u64()
        self.operator_quanta, self.wfn_cinfos = [], []
        for _ in range(r.i32()):
            self.operator_quanta.append(r.u64())
            self.wfn_cinfos.append(r.cinfo() if r.u8() else None)
        if r.p != len(r.b):
            raise FormatError("%d bytes after the EFFPROB1 record"
                              % (len(r.b) - r.p))


def load_eff_problem(path):
    """Parse an EFFPROB1 file (block2's EFF-PROB.FORW.i / EFF-PROB.BACKW.i)."""
    with open(path, "rb") as f:
        return EffProblemFile(f.read())


def load_operator_tensor(path):
    """Parse a file written by block2's OperatorTensor::save_data(ofs)
    (operator_tensor.hpp:122-148, plain storage, not pointer_only): a dict
    with the symbolic markers lmat / rmat and ops, a list of (key
    expression, dict(type, info, factor, data)).  block2 writes operators
    that were never allocated (those of a delayed tensor) with no
    elements."""
    with open(path, "rb") as f:
        return tensor_from_bytes(f.read())


def tensor_from_bytes(buf):
    r = _Reader(buf)
    t = r.data_tensor()
    if r.p != len(r.b):
        raise FormatError("%d bytes after the operator tensor"
                          % (len(r.b) - r.p))
    return t


def attach_operator_data(pb, left, right, dleft=None, dright=None):
    """Fill the operator buffers (and factors) of a Problem read by
    to_problem from parsed operator tensors (load_operator_tensor): the
    environment tensors `left` / `right` and, for a delayed side, the
    tensors of its sub-operators `dleft` / `dright`.  Operators are found
    by key; each must carry exactly the elements its info describes."""
    keys = pb.op_keys
    bufs = [np.zeros(pb.data_len[0]), np.zeros(pb.data_len[1])]
    lists = [(pb.lops, pb.linfo, False, keys[0], left, 0),
             (pb.rops, pb.rinfo, False, keys[1], right, 1)]
    if pb.delayed_side:
        if dleft is None or dright is None:
            raise FormatError("a delayed problem needs the tensors of the "
                              "delayed side's sub-operators")
        side = pb.delayed_side - 1
        lists += [(pb.dlops, pb.dlinfo, True, keys[2], dleft, side),
                  (pb.drops, pb.drinfo, True, keys[3], dright, side)]
    for ops, infos, own, kl, tensor, side in lists:
        if tensor is None:
            raise FormatError("operator tensor missing")
        by = {}
        for k, m in tensor["ops"]:
            by.setdefault(_key(k), m)
        for j, key in enumerate(kl):
            if ops.flags[j] & 1:
                continue
            m = by.get(key)
            n = infos.total(j if own else int(ops.info[j]))
            if m is None or m["data"] is None or len(m["data"]) != n:
                raise FormatError(
                    "operator %r has no data of %d elements in its tensor "
                    "(an operator formed from a delayed expression must be "
                    "formed by the caller)" % (key, n))
            o = int(ops.off[j])
            bufs[side][o:o + n] = m["data"]
            ops.factor[j] = m["factor"]
    pb.lops.data, pb.rops.data = bufs
    return pb


def _key(e):
    if e is None or e[0] != "elem":
        raise FormatError("operator of a string is not an element")
    return e[1:]


def _info_set(infos):
    n = [i["n"] for i in infos]
    off = np.concatenate([[0], np.cumsum(n)[:-1]]) if n else []
    cat = (lambda k, dt: np.concatenate([i[k] for i in infos]).astype(dt)
           if infos else np.zeros(0, dt))
    return InfoSet([i["dq"] for i in infos], [i["fermion"] for i in infos],
                   n, off, cat("quanta", np.uint64), cat("nbra", np.uint32),
                   cat("nket", np.uint32), cat("shift", np.uint64))


def _total(info):
    if info["n"] == 0:
        return 0
    return int(info["shift"][-1]) + int(info["nbra"][-1]) * int(info["nket"][-1])


class _Side:
    """Operators of one tensor, numbered by first use."""

    def __init__(self, tensor):
        self.t = tensor
        self.index = {}
        for i, (k, _) in enumerate(tensor["ops"]):
            self.index.setdefault(_key(k), i)
        self.exprs = ({_key(k) for k, _, _ in tensor["exprs"]}
                      if tensor["delayed"] else set())
        self.ids, self.mats, self.formed = {}, [], []

    def get(self, e, prod):
        k = _key(e)
        if k in self.ids:
            return self.ids[k]
        if k not in self.index:
            raise FormatError("string operator missing from its tensor")
        self.ids[k] = len(self.mats)
        self.mats.append(self.t["ops"][self.index[k]][1])
        m = self.mats[-1]
        self.formed.append(not self.t["delayed"] or bool(m["total"]) or
                           (prod and k in self.exprs))
        return self.ids[k]


def _strings(f):
    mat = f.op["mat"]
    if mat is None or not mat["data"]:
        raise FormatError("effective operator without an H expression")
    e = mat["data"][0]
    if e[0] == "exprref":
        e = e[2]
    if e[0] == "sum":
        return e[1]
    if e[0] in ("prod", "sumprod"):
        return [e]
    raise FormatError("H expression of type %s" % e[0])


def to_problem(f):
    """The structure-only Problem of an EFFPROB1 file (Problem.lops.data and
    rops.data None, Problem.data_len the element counts of the two operator
    buffers, sub-operators of a delayed tensor included)."""
    if f.op.get("stacked_mat") is not None:
        raise FormatError("stacked (fused) expressions are not supported")
    lt, rt = f.op["lopt"], f.op["ropt"]
    if lt is None or rt is None:
        raise FormatError("effective operator without left/right tensors")
    if lt["delayed"] and rt["delayed"]:
        raise FormatError("both environment tensors delayed")
    ls, rs = _Side(lt), _Side(rt)
    dside = 1 if lt["delayed"] else 2 if rt["delayed"] else 0
    dt = lt if dside == 1 else rt if dside == 2 else None
    dls = _Side(dt["lopt"]) if dt else None
    drs = _Side(dt["ropt"]) if dt else None
    conj, ta, tb, tf = [], [], [], []
    kind, tda, tdb, tdc = [], [], [], []
    for s in _strings(f):
        prod = s[0] == "prod"
        if s[0] not in ("prod", "sumprod"):
            raise FormatError("H string of type %s" % s[0])
        conj.append(s[2])
        ta.append(ls.get(s[3], prod))
        tb.append(rs.get(s[4], prod))
        tf.append(s[1])
        if prod:
            kind.append(0), tda.append(-1), tdb.append(-1), tdc.append(0)
        else:
            ops, cj = s[6], s[7]
            if dt is None or len(ops) != 2:
                raise FormatError("SumProd string without a delayed tensor")
            kind.append(1)
            tda.append(dls.get(ops[0], False))
            tdb.append(drs.get(ops[1], False))
            tdc.append((cj[0] & 1) | ((cj[1] & 1) << 1))

    def infos_of(side):
        # one info per operator quantum, in quantum order (block2's
        # left/right_op_infos), taken from the tensor's operators
        by = {}
        for _, m in side.t["ops"]:
            by.setdefault(int(m["info"]["dq"]), m)
        dqs = sorted(by)
        return dqs, [by[q]["info"] for q in dqs], [by[q]["cinfo"] for q in dqs]

    def op_set(side, dqs, own=False):
        dq, info, off, flags = [], [], [], []
        n = 0
        for j, m in enumerate(side.mats):
            q = int(m["info"]["dq"])
            dq.append(q)
            info.append(j if own else dqs.index(q))
            off.append(n)
            flags.append(0 if side.formed[j] else 1)
            if side.formed[j]:
                n += _total(m["info"])
        return OpSet(dq, np.ones(len(dq)), info, off, None, flags), n

    ldq, linf, lci = infos_of(ls)
    rdq, rinf, rci = infos_of(rs)
    lops, ln = op_set(ls, ldq)
    rops, rn = op_set(rs, rdq)
    ket = _info_set([f.cmat["info"]])
    bra = _info_set([f.vmat["info"]])
    pb = Problem(f.sym, f.opdq, 0, ket, bra, _info_set(linf),
                 _info_set(rinf), lops, rops, conj, ta, tb, tf,
                 dict(isite=f.isite, forward=f.forward, source="EFFPROB1"))
    pb.term_kind = np.asarray(kind, np.uint8)
    pb.term_da = np.asarray(tda, np.int32)
    pb.term_db = np.asarray(tdb, np.int32)
    pb.term_dconj = np.asarray(tdc, np.uint8)
    if dside:
        for sd in (dls, drs):
            sd.formed = [True] * len(sd.mats)
        dlops, dln = op_set(dls, None, own=True)
        drops, drn = op_set(drs, None, own=True)
        base = ln if dside == 1 else rn
        dlops.off = dlops.off + np.uint64(base)
        drops.off = drops.off + np.uint64(base + dln)
        if dside == 1:
            ln += dln + drn
        else:
            rn += dln + drn
        pb.delayed_side = dside
        pb.dlinfo = _info_set([m["info"] for m in dls.mats])
        pb.drinfo = _info_set([m["info"] for m in drs.mats])
        pb.dlops, pb.drops = dlops, drops
        for c in (lci if dside == 1 else rci):
            if c is None:
                pb.tp.append(dict(n=np.full(5, -1, np.int32), nc=0,
                                  **{k: np.zeros(0, dt) for k, dt in (
                                      ("quanta", np.uint64), ("idx", np.uint32),
                                      ("ia", np.uint32), ("ib", np.uint32),
                                      ("ic", np.uint32), ("stride", np.uint64),
                                      ("factor", np.float64))}))
            else:
                pb.tp.append(dict(c, nc=int(c["nc"])))
    pb.data_len = {0: ln, 1: rn}
    order = lambda sd: sorted(sd.ids, key=sd.ids.get) if sd else []
    pb.op_keys = [order(ls), order(rs), order(dls), order(drs)]
    return pb
