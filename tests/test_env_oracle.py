"""Environment blocking and renormalisation on the CPU: the oracle
(oracle/env_oracle.py) against block2's own dumped tensors, and the
library's host-side plans against the oracle.

block2's sweep (reference harness, tests/golden/make_sweep_fixtures.py)
dumped at several sites of a forward and a backward sweep: the
environments an effective Hamiltonian was blocked from, the blocked
operators and the MPS tensors that renormalised them.  Checks:
  * the oracle's left / right contraction (tensor_product over
    initialize_tp) of the dumped environment and the MPO's site operators
    reproduces block2's blocked operators;
  * the oracle's rotation of the previous site's blocked operators by the
    dumped MPS tensor reproduces block2's next environment;
  * the library's tensor-product connection lists (b2g_contract_plan,
    host-only) equal the oracle's entry by entry;
  * the fused bases and block layouts of b2g.states match the dumps.
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import REPO  # noqa: F401  (sets sys.path)
import env_oracle as E
from b2g import _lib as L
from b2g import mpo as M
from b2g import states as Q
from sweep_common import (SWEEP_CASES, basis_of, case_dir, contraction_terms,
                          key, load_ops, load_rot)

ENV_SITES = []
for case in SWEEP_CASES:
    d = case_dir(case)
    if os.path.isdir(d):
        for e in sorted(os.listdir(d)):
            if e.startswith("env_s"):
                ENV_SITES.append((case, e))


def _site(case, env):
    sw, site = env[len("env_s"):].split("_")
    return int(sw), int(site)


@pytest.fixture(scope="module")
def mpos():
    return {c: M.load_mpo(os.path.join(case_dir(c), "mpo.bin"))
            for c in SWEEP_CASES if os.path.isdir(case_dir(c))}


def _block(case, env, mpos, left):
    sym = SWEEP_CASES[case]
    mpo = mpos[case]
    sw, c = _site(case, env)
    d = os.path.join(case_dir(case), env)
    site = c if left else c + 1
    envf = os.path.join(d, "lenv.bin" if left else "renv.bin")
    blk = load_ops(os.path.join(d, "lblk.bin" if left else "rblk.bin"))
    if not os.path.exists(envf):
        return None
    env_ops = load_ops(envf)
    site_ops = {key(k): dict(
        dq=int(m["info"]["dq"]), factor=float(m["factor"]),
        info=dict(quanta=[int(x) for x in m["info"]["quanta"]],
                  nbra=[int(x) for x in m["info"]["nbra"]],
                  nket=[int(x) for x in m["info"]["nket"]],
                  shift=[int(x) for x in m["info"]["shift"]],
                  dq=int(m["info"]["dq"])),
        data=np.asarray(m["data"], np.float64))
        for k, m in mpo["tensors"][site]["ops"]}
    ek, sk, bk = list(env_ops), list(site_ops), list(blk)
    terms, inters = contraction_terms(mpo, site, left, ek, sk, bk)
    sb = (list(map(int, mpo["basis"][site]["quanta"])),
          list(map(int, mpo["basis"][site]["n_states"])))
    eb = basis_of(env_ops, E, sym)
    fused = basis_of(blk, E, sym)
    # SumProd strings: the summed environment operators as intermediates
    # (OperatorFunctions::iadd into an operator of the first one's shape)
    env_list = list(env_ops.values())
    for j, (dq, parts) in enumerate(inters):
        inf = env_list[parts[0][0]]["info"]
        tmp = dict(dq=dq, factor=1.0, info=inf,
                   data=np.zeros(inf["total"]))
        for ia, cj, f in parts:
            E.iadd(sym, tmp, env_list[ia], f, cj)
        env_ops[("tmp", j)] = tmp
    return dict(sym=sym, env=env_ops, site=site_ops, blk=blk, terms=terms,
                sb=sb, eb=eb, fused=fused, ek=ek, sk=sk, bk=bk)


@pytest.mark.parametrize("case,env", ENV_SITES)
@pytest.mark.parametrize("left", [True, False], ids=["left", "right"])
def test_oracle_blocking_matches_block2(case, env, left, mpos):
    b = _block(case, env, mpos, left)
    if b is None:
        pytest.skip("no environment on this side at this site")
    c_ops = [dict(dq=op["dq"], info=op["info"]) for op in b["blk"].values()]
    if left:
        E.contract(b["sym"], list(b["env"].values()),
                   list(b["site"].values()), c_ops, b["terms"], b["eb"],
                   b["eb"], b["sb"], b["sb"], b["fused"], b["fused"])
    else:
        E.contract(b["sym"], list(b["site"].values()),
                   list(b["env"].values()), c_ops, b["terms"], b["sb"],
                   b["sb"], b["eb"], b["eb"], b["fused"], b["fused"])
    worst = 0.0
    for got, ref in zip(c_ops, b["blk"].values()):
        r = ref["data"] * ref["factor"]
        scale = max(np.abs(r).max() if r.size else 0.0, 1.0)
        worst = max(worst, (np.abs(got["data"] - r).max() if r.size else 0.0)
                    / scale)
    assert worst < 1e-13, worst


def _rotation_pairs():
    """(case, env of the blocked operators, env of the next environment,
    left) where block2 renormalised the first site's blocked operators into
    the second site's environment."""
    out = []
    for case, env in ENV_SITES:
        sw, c = _site(case, env)
        if sw % 2 == 0 and (case, "env_s%d_%d" % (sw, c + 1)) in ENV_SITES:
            out.append((case, env, "env_s%d_%d" % (sw, c + 1), True))
        if sw % 2 == 1 and (case, "env_s%d_%d" % (sw, c - 1)) in ENV_SITES:
            out.append((case, env, "env_s%d_%d" % (sw, c - 1), False))
    return out


@pytest.mark.parametrize("case,src,dst,left", _rotation_pairs())
def test_oracle_rotation_matches_block2(case, src, dst, left):
    sym = SWEEP_CASES[case]
    blk = load_ops(os.path.join(case_dir(case), src,
                                "lblk.bin" if left else "rblk.bin"))
    nxt = os.path.join(case_dir(case), dst)
    env = load_ops(os.path.join(nxt, "lenv.bin" if left else "renv.bin"))
    rot = load_rot(os.path.join(nxt, "mps_left.bin" if left else
                                "mps_right.bin"))
    n = 0
    for k, ref in env.items():
        if k not in blk or ref["data"] is None or ref["data"].size == 0:
            # formed by the schemer's transform, or released by it
            continue
        c = dict(dq=ref["dq"], info=ref["info"])
        E.rotate(sym, blk[k], rot, c, not left)
        r = ref["data"] * ref["factor"]
        assert np.abs(c["data"] - r).max() <= 1e-12 * max(np.abs(r).max(),
                                                           1.0), k
        n += 1
    assert n > 0


@pytest.mark.parametrize("case,env", ENV_SITES)
@pytest.mark.parametrize("left", [True, False], ids=["left", "right"])
def test_host_tp_connections_match_oracle(case, env, left, mpos):
    b = _block(case, env, mpos, left)
    if b is None:
        pytest.skip("no environment on this side at this site")
    sym = b["sym"]
    keep = []

    def side(ops):
        lays = {}
        dqs = []
        for op in ops.values():
            i = op["info"]
            lays[op["dq"]] = Q.Layout(i["quanta"], i["nbra"], i["nket"],
                                      op["dq"], Q.q_is_fermion(op["dq"]))
            dqs.append(op["dq"])
        idq = sorted(lays)
        arr = (L.b2g_sparse_info * len(idq))(*[L.b2g_sparse_info(
            q, int(lays[q].fermion), len(lays[q]),
            L.ptr(lays[q].quanta, C.c_uint64),
            L.ptr(lays[q].nbra, C.c_uint32),
            L.ptr(lays[q].nket, C.c_uint32),
            L.ptr(lays[q].shift, C.c_uint64)) for q in idq])
        off, t = [], 0
        for q in dqs:
            off.append(t)
            t += lays[q].total
        ops_c = (L.b2g_operator * len(dqs))(*[L.b2g_operator(
            dqs[i], 1.0, idq.index(dqs[i]), 0, off[i])
            for i in range(len(dqs))])
        dqa = np.array(idq, np.uint64)
        keep.extend([arr, ops_c, dqa, lays])
        return len(idq), L.ptr(dqa, C.c_uint64), arr, len(dqs), ops_c, idq, \
            lays

    def si(b_):
        q = np.array(b_[0], np.uint64)
        n = np.array(b_[1], np.uint32)
        keep.extend([q, n])
        return L.b2g_state_info(len(q), 0, L.ptr(q, C.c_uint64),
                                L.ptr(n, C.c_uint32))

    A, Bs = (b["env"], b["site"]) if left else (b["site"], b["env"])
    ab, bb = (b["eb"], b["sb"]) if left else (b["sb"], b["eb"])
    sa, sb_, sc = side(A), side(Bs), side(b["blk"])
    c = L.b2g_contract()
    c.symmetry = sym
    c.bra_a = c.ket_a = si(ab)
    c.bra_b = c.ket_b = si(bb)
    c.bra = c.ket = si(b["fused"])
    c.n_a_infos, c.a_info_dq, c.a_infos, c.n_a_ops, c.a_ops = sa[:5]
    c.n_b_infos, c.b_info_dq, c.b_infos, c.n_b_ops, c.b_ops = sb_[:5]
    c.n_c_infos, c.c_info_dq, c.c_infos, c.n_c_ops, c.c_ops = sc[:5]
    tm = (L.b2g_contract_term * len(b["terms"]))(
        *[L.b2g_contract_term(*t) for t in b["terms"]])
    c.n_terms, c.terms = len(b["terms"]), tm
    lib = L.lib()
    h = C.c_void_p()
    L.check(lib.b2g_contract_plan(None, C.byref(c), L.FLAG_HOST_ONLY,
                                  C.byref(h)))
    try:
        a_list, b_list = list(A.values()), list(Bs.values())
        blk = list(b["blk"].values())
        ainf = {op["dq"]: op["info"] for op in a_list}
        binf = {op["dq"]: op["info"] for op in b_list}
        for ii, q in enumerate(sc[5]):
            pairs = [(t[3], a_list[t[1]]["dq"], b_list[t[2]]["dq"])
                     for t in b["terms"] if blk[t[0]]["dq"] == q]
            subl = E.sub_labels(sym, q, pairs)
            cinfo = next(op["info"] for op in blk if op["dq"] == q)
            ref = E.tp_connections(sym, q, subl, b["fused"], b["fused"], ab,
                                   bb, ab, bb, ainf, binf, cinfo)
            exp = [x for cl in subl for x in ref[cl]]
            n5 = (C.c_int32 * 5)()
            ncl, nc = C.c_uint64(), C.c_uint64()
            L.check(lib.b2g_winplan_connections(
                h, ii, n5, C.byref(ncl), None, None, C.byref(nc), None, None,
                None, None, None))
            m = nc.value
            arrs = [np.zeros(m, np.uint32) for _ in range(3)]
            st, fa = np.zeros(m, np.uint64), np.zeros(m)
            qs, idx = np.zeros(ncl.value, np.uint64), \
                np.zeros(ncl.value, np.uint32)
            L.check(lib.b2g_winplan_connections(
                h, ii, n5, C.byref(ncl), L.ptr(qs, C.c_uint64),
                L.ptr(idx, C.c_uint32), C.byref(nc),
                *[L.ptr(a, C.c_uint32) for a in arrs], L.ptr(st, C.c_uint64),
                L.ptr(fa, C.c_double)))
            assert [tuple(int(x) for x in s) for s in subl] == \
                [(n5_c, int(qq)) for n5_c, qq in zip(
                    [next(k for k in range(4) if n5[k] <= j < n5[k + 1])
                     for j in range(ncl.value)], qs)]
            assert m == len(exp)
            for k in range(m):
                e = exp[k]
                assert (int(arrs[0][k]), int(arrs[1][k]), int(arrs[2][k]),
                        int(st[k])) == e[:4]
                assert abs(fa[k] - e[4]) <= 1e-14 * abs(e[4])
    finally:
        lib.b2g_winplan_destroy(h)


@pytest.mark.parametrize("case", [c for c in SWEEP_CASES
                                  if os.path.isdir(case_dir(c))])
def test_fused_layouts_match_block2(case):
    """b2g.states' StateInfo fusion and SparseMatrixInfo::initialize give
    the layouts of block2's dumped blocked operators and MPS tensors."""
    sym = SWEEP_CASES[case]
    mps = M.load_mps(case_dir(case))
    info = mps["info"]
    basis = [Q.StateInfo.from_dict(b) for b in info["basis"]]
    ldf = [Q.StateInfo.from_dict(b) for b in info["left_dims_fci"]]
    rdf = [Q.StateInfo.from_dict(b) for b in info["right_dims_fci"]]
    n = mps["n_sites"]
    vac = Q.StateInfo([int(info["vacuum"])], [1])
    right = [None] * (n + 1)
    right[n] = vac
    for i in range(2, n):
        right[i] = Q.extract_state_info(
            sym, Q.Layout.from_dict(mps["tensors"][i]["info"]), False)
    # the starting two-site tensor
    fl = Q.tensor_product(sym, vac, basis[0], ldf[1])
    fr = Q.tensor_product(sym, basis[1], right[2], rdf[1])
    lay = Q.initialize(sym, fl, fr, int(info["target"]), False, True)
    assert lay.same(Q.Layout.from_dict(mps["tensors"][0]["info"]))
    # right-canonical tensors: [right_dims[i] x (basis[i] (x) right[i+1])]
    for i in range(2, n):
        f = Q.tensor_product(sym, basis[i], right[i + 1], rdf[i])
        lay = Q.initialize(sym, right[i], f, int(info["vacuum"]), False)
        assert lay.same(Q.Layout.from_dict(mps["tensors"][i]["info"])), i
    # blocked operators of the dumped sites
    for case_, env in ENV_SITES:
        if case_ != case:
            continue
        blk = load_ops(os.path.join(case_dir(case), env, "lblk.bin"))
        fused = Q.StateInfo(*basis_of(blk, E, sym))
        for op in blk.values():
            lay = Q.initialize(sym, fused, fused, op["dq"],
                               Q.q_is_fermion(op["dq"]))
            i = op["info"]
            assert lay.same(Q.Layout(i["quanta"], i["nbra"], i["nket"],
                                     op["dq"])), (env, op["dq"])
