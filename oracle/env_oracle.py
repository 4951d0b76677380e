"""CPU oracle of environment blocking and renormalisation (TEST
INFRASTRUCTURE ONLY: imported by tests/ and nothing in the product).

A numpy / pure-Python restatement of block2's steps either side of H*psi in
a two-site sweep, written for small problems and checked against block2's
own outputs dumped by the reference harness (oracle/ref_harness.cpp):

  fuse                  StateInfo::tensor_product(a, b, cref)
                        src/core/state_info.hpp:231-254
  connection            StateInfo::get_connection_info  :283-312
  op_info               SparseMatrixInfo::initialize
                        src/core/sparse_matrix.hpp:727-757
  tp_connections        ConnectionInfo::initialize_tp   :294-422
  contract              TensorFunctions::left_contract / right_contract
                        src/core/tensor_functions.hpp:2842-2986, the Prod
                        case of tensor_product :2184-2288,
                        OperatorFunctions::tensor_product
                        src/core/operator_functions.hpp:672-711,
                        GMatrixFunctions::tensor_product
                        src/core/matrix_functions.hpp:1269-
  rotate                OperatorFunctions::tensor_rotate
                        src/core/operator_functions.hpp:175-211
                        (GMatrixFunctions::rotate, matrix_functions.hpp:998-1009)
  iadd                  OperatorFunctions::iadd  operator_functions.hpp:134-174
  swap_to_fused_left/right  SparseMatrix::swap_to_fused_left / right
                        src/core/sparse_matrix.hpp:1823-1968
  contract_wfn          SparseMatrix::contract  sparse_matrix.hpp:1777-1801

Labels are block2's packed SZ / SU2 words; the spin algebra is evaluated
with the oracle's own 6j / 9j (b2g_oracle.sixj / ninej, Racah's formula).
An operator is a dict: dq, factor, info (quanta, nbra, nket, shift lists),
data (numpy array).
"""
import math

import numpy as np

import b2g_oracle as O

SZ, SU2 = 0, 1
INVALID = (1 << 64) - 1


# ------------------------------------------------------------ label algebra
def _s16(v):
    return v - 0x10000 if v & 0x8000 else v


def n_of(q):
    return _s16((q >> 48) & 0xFFFF)


def pg_of(q):
    return q & 0xFFFF


def twos(s, q):
    v = (q >> 16) & 0xFFFF
    return _s16(v) if s == SZ else v


def twos_low(q):
    return (q >> 32) & 0xFFFF


def make(s, n, tl, t, pg):
    w = ((n & 0xFFFF) << 48) | ((t & 0xFFFF) << 16) | (pg & 0xFFFF)
    return w | ((tl & 0xFFFF) << 32) if s == SU2 else w


def neg(s, q):
    if s == SZ:
        return make(s, -n_of(q), 0, -twos(s, q), pg_of(q))
    return make(s, -n_of(q), twos_low(q), twos(s, q), pg_of(q))


def members(s, a, b):
    """The labels of a + b: one for SZ, the spin range for SU2."""
    n, pg = n_of(a) + n_of(b), pg_of(a) ^ pg_of(b)
    if s == SZ:
        return [make(s, n, 0, twos(s, a) + twos(s, b), pg)]
    ta, tb = twos(s, a), twos(s, b)
    return [make(s, n, t, t, pg) for t in range(abs(ta - tb), ta + tb + 1, 2)]


def ket_of(s, q):
    if s == SZ:
        return q
    return make(s, n_of(q), twos(s, q), twos(s, q), pg_of(q))


def bra_of(s, q, dq):
    if s == SZ:
        return members(s, q, dq)[0]
    t = twos_low(q)
    return make(s, n_of(q) + n_of(dq), t, t, pg_of(q) ^ pg_of(dq))


def combine(s, dq, bra, ket):
    """Composite label of <bra| op(dq) |ket> or INVALID."""
    if s == SZ:
        return ket if members(s, ket, dq)[0] == bra else INVALID
    tb, tk, td = twos(s, bra), twos(s, ket), twos(s, dq)
    if n_of(bra) != n_of(ket) + n_of(dq) or \
            pg_of(bra) != pg_of(ket) ^ pg_of(dq):
        return INVALID
    if (tb + tk + td) & 1 or not abs(tk - td) <= tb <= tk + td:
        return INVALID
    return make(s, n_of(ket), tb, tk, pg_of(ket))


def mult(s, q):
    return 1 if s == SZ else twos(s, q) + 1


def is_fermion(q):
    return bool((q >> 16) & 1)


def transpose_cg(s, td, tl, tr):
    if s == SZ:
        return 1.0
    r = math.sqrt((tr + 1) / (tl + 1))
    return -r if (td + tl - tr) & 2 else r


def racah(s, a, b, c, d, e, f):
    """CG::racah: (-1)^((a+b+c+d)/2) {a b e; d c f} on doubled spins."""
    if s == SZ:
        return 1.0
    ph = -1.0 if (a + b + c + d) & 2 else 1.0
    return ph * O.sixj(a, b, e, d, c, f)


# --------------------------------------------------------- StateInfo, info
def find(qs, q):
    i = int(np.searchsorted(np.asarray(qs, np.uint64), np.uint64(q)))
    return i if i < len(qs) and int(qs[i]) == q else -1


def fuse(s, a, b, cref):
    """StateInfo::tensor_product(a, b, cref); a, b, cref: (quanta, counts)."""
    n = [0] * len(cref[0])
    for qa, na in zip(*a):
        for qb, nb in zip(*b):
            for q in members(s, int(qa), int(qb)):
                i = find(cref[0], q)
                if i >= 0:
                    n[i] += int(na) * int(nb)
    keep = [i for i in range(len(n)) if n[i]]
    return ([int(cref[0][i]) for i in keep], [n[i] for i in keep])


def connection(s, a, b, c):
    """StateInfo::get_connection_info: acc, pairs."""
    mp = {}
    for i, qa in enumerate(a[0]):
        for j, qb in enumerate(b[0]):
            for q in members(s, int(qa), int(qb)):
                mp.setdefault(q, []).append((i, j))
    acc, ij = [0], []
    for q in c[0]:
        ij += mp[int(q)]
        acc.append(len(ij))
    return acc, ij


def op_info(s, bra, ket, dq, wfn=False):
    """SparseMatrixInfo::initialize(bra, ket, dq, ..., wfn)."""
    qs = []
    for q in ket[0]:
        q = neg(s, int(q)) if wfn else int(q)
        for b in members(s, dq, q):
            if find(bra[0], b) >= 0:
                qs.append(combine(s, dq, b, q))
    qs.sort()
    nb, nk = [], []
    for x in qs:
        k = ket_of(s, x)
        nk.append(int(ket[1][find(ket[0], neg(s, k) if wfn else k)]))
        nb.append(int(bra[1][find(bra[0], bra_of(s, x, dq))]))
    shift, t = [], 0
    for a, b in zip(nb, nk):
        shift.append(t)
        t += a * b
    return dict(quanta=qs, nbra=nb, nket=nk, shift=shift, dq=dq, total=t)


def block(op, i):
    inf = op["info"]
    o, m, n = int(inf["shift"][i]), int(inf["nbra"][i]), int(inf["nket"][i])
    return op["data"][o:o + m * n].reshape(m, n)


def info_find(inf, q):
    return find(inf["quanta"], q)


# ------------------------------------------------------- tensor products
def tp_connections(s, cdq, subdq, bra, ket, bra_a, bra_b, ket_a, ket_b,
                   ainfos, binfos, cinfo):
    """ConnectionInfo::initialize_tp: per class (conj, label) the entries
    (ia, ib, ic, stride, factor).  ainfos / binfos: dict dq -> info."""
    bc = connection(s, bra_a, bra_b, bra)
    kc = connection(s, ket_a, ket_b, ket)
    out = {}
    for conj, sq in subdq:
        cja, cjb = conj & 1, (conj >> 1) & 1
        adq = neg(s, bra_of(s, sq, cdq)) if cja else bra_of(s, sq, cdq)
        bdq = ket_of(s, sq) if cjb else neg(s, ket_of(s, sq))
        ent = []
        out[(conj, sq)] = ent
        if adq not in ainfos or bdq not in binfos:
            continue
        ai, bi = ainfos[adq], binfos[bdq]
        for ic, cqq in enumerate(cinfo["quanta"]):
            cq, cqp = bra_of(s, cqq, cdq), ket_of(s, cqq)
            ib, ik = find(bra[0], cq), find(ket[0], cqp)
            bstr = 0
            for kb in range(bc[0][ib], bc[0][ib + 1]):
                jba, jbb = bc[1][kb]
                kstr = 0
                for kk in range(kc[0][ik], kc[0][ik + 1]):
                    jka, jkb = kc[1][kk]
                    aq, aqp = int(bra_a[0][jba]), int(ket_a[0][jka])
                    bq, bqp = int(bra_b[0][jbb]), int(ket_b[0][jkb])
                    qa = combine(s, adq, aqp, aq) if cja else \
                        combine(s, adq, aq, aqp)
                    qb = combine(s, bdq, bqp, bq) if cjb else \
                        combine(s, bdq, bq, bqp)
                    if qa != INVALID and qb != INVALID:
                        xa, xb = info_find(ai, qa), info_find(bi, qb)
                        if xa >= 0 and xb >= 0:
                            f = 1.0
                            if s == SU2:
                                f = math.sqrt(mult(s, cqp) * mult(s, cdq) *
                                              mult(s, aq) * mult(s, bq)) * \
                                    O.ninej(twos(s, aqp), twos(s, bqp),
                                            twos(s, cqp), twos(s, adq),
                                            twos(s, bdq), twos(s, cdq),
                                            twos(s, aq), twos(s, bq),
                                            twos(s, cq))
                            if is_fermion(bdq) and is_fermion(aqp):
                                f = -f
                            if cja:
                                f *= transpose_cg(s, twos(s, adq),
                                                  twos(s, aq), twos(s, aqp))
                            if cjb:
                                f *= transpose_cg(s, twos(s, bdq),
                                                  twos(s, bq), twos(s, bqp))
                            if abs(f) >= 1e-20:
                                ent.append((xa, xb, ic,
                                            bstr * int(cinfo["nket"][ic]) +
                                            kstr, f))
                    kstr += int(ket_a[1][jka]) * int(ket_b[1][jkb])
                bstr += int(bra_a[1][jba]) * int(bra_b[1][jbb])
    return out


def sub_labels(s, cdq, pairs):
    """Partition::get_uniq_sub_labels for one operator quantum: the
    (conj, cdq.combine(bra, ket)) of its strings plus the vacuum class;
    pairs: (conj, adq, bdq)."""
    out = set()
    for conj, adq, bdq in pairs:
        bra = neg(s, adq) if conj & 1 else adq
        ket = bdq if conj & 2 else neg(s, bdq)
        p = combine(s, cdq, bra, ket)
        assert p != INVALID
        out.add((conj, p))
    lo = members(s, cdq, neg(s, cdq))[0]
    p0 = combine(s, cdq, 0, lo)
    if p0 != INVALID:
        out.add((0, p0))
    return sorted(out)


def kron_into(c, a, conja, b, conjb, scale):
    """GMatrixFunctions::tensor_product: c += scale * op(a) (x) op(b)."""
    oa = a.T if conja else a
    ob = b.T if conjb else b
    c += scale * np.kron(oa, ob)


def contract(s, A, B, c_ops, terms, bra_a, ket_a, bra_b, ket_b, bra, ket):
    """The operators c_ops (dicts with dq, info; data filled here) as sums
    of tensor products: terms (c, a, b, conj, factor) over A[a] (x) B[b]."""
    ainfos = {op["dq"]: op["info"] for op in A}
    binfos = {op["dq"]: op["info"] for op in B}
    by_dq = {}
    for c, a, b, cj, f in terms:
        by_dq.setdefault(c_ops[c]["dq"], []).append(
            (cj, A[a]["dq"], B[b]["dq"]))
    for op in c_ops:
        by_dq.setdefault(op["dq"], [])
    tps = {}
    for dq, pairs in by_dq.items():
        inf = next(o["info"] for o in c_ops if o["dq"] == dq)
        tps[dq] = tp_connections(s, dq, sub_labels(s, dq, pairs), bra, ket,
                                 bra_a, bra_b, ket_a, ket_b, ainfos, binfos,
                                 inf)
    for op in c_ops:
        op["data"] = np.zeros(op["info"]["total"])
        op["factor"] = 1.0
    for c, a, b, cj, f in terms:
        co, ao, bo = c_ops[c], A[a], B[b]
        scale = f * ao["factor"] * bo["factor"]
        if abs(scale) < 1e-20:
            continue
        abdq = combine(s, co["dq"], neg(s, ao["dq"]) if cj & 1 else ao["dq"],
                       bo["dq"] if cj & 2 else neg(s, bo["dq"]))
        for xa, xb, ic, stride, fac in tps[co["dq"]][(cj, abdq)]:
            blk = block(co, ic)
            av, bv = block(ao, xa), block(bo, xb)
            m = (av.shape[1] if cj & 1 else av.shape[0]) * \
                (bv.shape[1] if cj & 2 else bv.shape[0])
            n = (av.shape[0] if cj & 1 else av.shape[1]) * \
                (bv.shape[0] if cj & 2 else bv.shape[1])
            r0, c0 = divmod(stride, blk.shape[1])
            kron_into(blk[r0:r0 + m, c0:c0 + n], av, cj & 1, bv, cj & 2,
                      scale * fac)
    return c_ops


def rotate(s, a, rot, c, trans):
    """OperatorFunctions::tensor_rotate: c[ic] += a.factor *
    op(rot_bra) a[ia] op(rot_ket) with rot_bra = rot_ket = rot (trans False:
    rot^T a rot; True: rot a rot^T)."""
    c["data"] = np.zeros(c["info"]["total"])
    for ic, q in enumerate(c["info"]["quanta"]):
        ia = info_find(a["info"], q)
        if ia < 0:
            continue
        cq, cqp = bra_of(s, q, c["dq"]), ket_of(s, q)
        ib, ik = info_find(rot["info"], cq), info_find(rot["info"], cqp)
        if ib < 0 or ik < 0:
            continue
        rb, rk, av = block(rot, ib), block(rot, ik), block(a, ia)
        blk = block(c, ic)
        if trans:
            blk += a["factor"] * (rb @ av @ rk.T)
        else:
            blk += a["factor"] * (rb.T @ av @ rk)
    return c


def iadd(s, c, a, scale, conj):
    """OperatorFunctions::iadd: c += scale * op(a) (transposed with block2's
    transpose factor when conj)."""
    for ic, q in enumerate(c["info"]["quanta"]):
        bra, ket = bra_of(s, q, c["dq"]), ket_of(s, q)
        aq = combine(s, a["dq"], ket, bra) if conj else \
            combine(s, a["dq"], bra, ket)
        if aq == INVALID:
            continue
        ia = info_find(a["info"], aq)
        if ia < 0:
            continue
        f = scale * a["factor"]
        if conj:
            f *= transpose_cg(s, twos(s, a["dq"]), twos(s, bra), twos(s, ket))
        av = block(a, ia)
        block(c, ic)[...] += f * (av.T if conj else av)


def contract_wfn(s, lmat, l_is_wfn, rmat, out_info):
    """SparseMatrix::contract: wfn = wfn x rot (l_is_wfn) or rot x wfn."""
    out = dict(info=out_info, dq=out_info["dq"],
               data=np.zeros(out_info["total"]))
    for i, q in enumerate(out_info["quanta"]):
        if l_is_wfn:
            il = info_find(lmat["info"], q)
            ir = info_find(rmat["info"], neg(s, ket_of(s, q)))
        else:
            il = info_find(lmat["info"], bra_of(s, q, out_info["dq"]))
            ir = info_find(rmat["info"], q)
        if il >= 0 and ir >= 0:
            block(out, i)[...] = block(lmat, il) @ block(rmat, ir)
    return out


def swap_to_fused_left(s, wfn, l, m, r, old_fused, new_fused, new_info):
    """SparseMatrix::swap_to_fused_left: [l x (m r)] -> [(l m) x r]."""
    dq = new_info["dq"]
    oc = connection(s, m, r, old_fused)
    nc = connection(s, l, m, new_fused)
    mp = {}
    oi = wfn["info"]
    for i, q in enumerate(oi["quanta"]):
        bra, ket = bra_of(s, q, oi["dq"]), neg(s, ket_of(s, q))
        ib, ik = find(l[0], bra), find(old_fused[0], ket)
        p = int(oi["shift"][i])
        for kk in range(oc[0][ik], oc[0][ik + 1]):
            ka, kb = oc[1][kk]
            mp.setdefault((ib, ka), {}).setdefault(kb, []).append((p, ik))
            p += int(m[1][ka]) * int(r[1][kb])
    out = np.zeros(new_info["total"])
    for i, q in enumerate(new_info["quanta"]):
        bra, ket = bra_of(s, q, dq), neg(s, ket_of(s, q))
        ib, ik = find(new_fused[0], bra), find(r[0], ket)
        ptr = int(new_info["shift"][i])
        for bb in range(nc[0][ib], nc[0][ib + 1]):
            ba, bm = nc[1][bb]
            lp = int(m[1][bm]) * int(r[1][ik])
            for p, iko in mp.get((ba, bm), {}).get(ik, []):
                ket_mr = int(old_fused[0][iko])
                f = racah(s, twos(s, int(l[0][ba])), twos(s, int(m[0][bm])),
                          twos(s, dq), twos(s, ket), twos(s, bra),
                          twos(s, ket_mr)) * \
                    math.sqrt(mult(s, bra) * mult(s, ket_mr))
                ld = int(old_fused[1][iko])
                for j in range(int(l[1][ba])):
                    out[ptr + j * lp:ptr + (j + 1) * lp] += \
                        f * wfn["data"][p + j * ld:p + j * ld + lp]
            ptr += int(l[1][ba]) * lp
    return dict(info=new_info, dq=dq, data=out)


def swap_to_fused_right(s, wfn, l, m, r, old_fused, new_fused, new_info):
    """SparseMatrix::swap_to_fused_right: [(l m) x r] -> [l x (m r)]."""
    dq = new_info["dq"]
    oc = connection(s, l, m, old_fused)
    nc = connection(s, m, r, new_fused)
    mp = {}
    oi = wfn["info"]
    for i, q in enumerate(oi["quanta"]):
        bra, ket = bra_of(s, q, oi["dq"]), neg(s, ket_of(s, q))
        ib, ik = find(old_fused[0], bra), find(r[0], ket)
        p = int(oi["shift"][i])
        for bb in range(oc[0][ib], oc[0][ib + 1]):
            ba, bm = oc[1][bb]
            lp = int(m[1][bm]) * int(r[1][ik])
            mp.setdefault((bm, ik), {}).setdefault(ba, []).append(
                (p, lp, ib))
            p += int(l[1][ba]) * lp
    out = np.zeros(new_info["total"])
    for i, q in enumerate(new_info["quanta"]):
        bra, ket = bra_of(s, q, dq), neg(s, ket_of(s, q))
        ib, ik = find(l[0], bra), find(new_fused[0], ket)
        ptr = int(new_info["shift"][i])
        lp = int(new_fused[1][ik])
        for kk in range(nc[0][ik], nc[0][ik + 1]):
            ka, kb = nc[1][kk]
            for p, w, ibo in mp.get((ka, kb), {}).get(ib, []):
                bra_lm = int(old_fused[0][ibo])
                f = racah(s, twos(s, int(r[0][kb])), twos(s, int(m[0][ka])),
                          twos(s, dq), twos(s, bra), twos(s, ket),
                          twos(s, bra_lm)) * \
                    math.sqrt(mult(s, ket) * mult(s, bra_lm))
                for j in range(int(l[1][ib])):
                    out[ptr + j * lp:ptr + j * lp + w] += \
                        f * wfn["data"][p + j * w:p + (j + 1) * w]
            ptr += int(m[1][ka]) * int(r[1][kb])
    return dict(info=new_info, dq=dq, data=out)
