"""Shared helpers of the environment and sweep tests: block2's dumped
operator tensors (OperatorTensor::save_data files in the sweep fixtures)
as oracle operators, and the contraction strings of an MPO site."""
import os

import numpy as np

from conftest import REPO

from b2g import effprob as EP
from b2g import mpo as M

GOLDEN = os.path.join(REPO, "tests", "golden")
SWEEP_CASES = {"hub8_sz": 0, "hub8_su2": 1, "hub8_sz_quiet": 0,
               "hub8_su2_quiet": 1, "hub16_sz": 0, "n2_su2": 1,
               "c2_su2": 1, "hub8_sz_pert": 0, "hub8_su2_pert": 1,
               "n2_su2_pert": 1, "hub16_sz_pert": 0}


# the truncation cases dumped at the sweep sites (split.npz): wavefunction
# noise trace_right (sW), density-matrix noise trace_left (sD), and
# density-matrix noise with every state kept, trace_right (sE) and
# trace_left (sF); noise 1e-6 from block2's generator seeded with 4321
SWEEP_SPLIT_CASES = ["sW", "sD", "sE", "sF"]


def case_dir(case):
    return os.path.join(GOLDEN, "sweep_" + case)


def sweep_split_fixtures():
    out = []
    for case in SWEEP_CASES:
        d = case_dir(case)
        if not os.path.isdir(d):
            continue
        for e in sorted(os.listdir(d)):
            if os.path.exists(os.path.join(d, e, "split.npz")):
                out.append((case, e))
    return out


def load_sweep_split(case, env):
    """(arrays, symmetry 0 = SZ / 1 = SU(2), wavefunction layout dict) of a
    sweep site's truncation dump"""
    z = np.load(os.path.join(case_dir(case), env, "split.npz"))
    info = dict(quanta=z["ket_quanta"], nbra=z["ket_nbra"],
                nket=z["ket_nket"], shift=z["ket_shift"],
                dq=int(z["ket_dq"][0]))
    return z, SWEEP_CASES[case], info


def key(e):
    return (e[1], e[2], e[4])


def load_ops(path):
    """{key: oracle operator} of a dumped OperatorTensor, in file order."""
    t = EP.load_operator_tensor(path)
    out = {}
    for k, m in t["ops"]:
        inf = m["info"]
        q = [int(x) for x in inf["quanta"]]
        nb = [int(x) for x in inf["nbra"]]
        nk = [int(x) for x in inf["nket"]]
        sh = [int(x) for x in inf["shift"]]
        out[key(k)] = dict(dq=int(inf["dq"]), factor=float(m["factor"]),
                           info=dict(quanta=q, nbra=nb, nket=nk, shift=sh,
                                     dq=int(inf["dq"]),
                                     total=sum(a * b for a, b in zip(nb, nk))),
                           data=(None if m["data"] is None else
                                 np.asarray(m["data"], np.float64)))
    return out


def load_rot(path):
    m = M.load_sparse_matrix(path)
    i = m["info"]
    return dict(dq=int(i["dq"]),
                info=dict(quanta=[int(x) for x in i["quanta"]],
                          nbra=[int(x) for x in i["nbra"]],
                          nket=[int(x) for x in i["nket"]],
                          shift=[int(x) for x in i["shift"]], dq=int(i["dq"])),
                data=m["data"] * m["factor"])


def basis_of(ops, env_oracle, sym, bra=True):
    """Row (bra) or column basis of an operator set, from its blocks."""
    qs = {}
    for op in ops.values():
        inf = op["info"]
        for q, nb, nk in zip(inf["quanta"], inf["nbra"], inf["nket"]):
            if bra:
                qs[env_oracle.bra_of(sym, q, op["dq"])] = nb
            else:
                qs[env_oracle.ket_of(sym, q)] = nk
    k = sorted(qs)
    return (k, [qs[x] for x in k])


def strings(e):
    """(factor, conj, a, b) of a sum of products (None for other forms)."""
    if e[0] == "exprref":
        e = e[2]
    if e[0] == "zero":
        return []
    if e[0] == "prod":
        return [(e[1], e[2], e[3], e[4])]
    if e[0] == "sum" and all(x[0] == "prod" for x in e[1]):
        return [(x[1], x[2], x[3], x[4]) for x in e[1]]
    return None


def _all_strings(e):
    if e[0] == "exprref":
        e = e[2]
    if e[0] == "zero":
        return []
    if e[0] in ("prod", "sumprod"):
        return [e]
    if e[0] == "sum":
        return list(e[1])
    raise ValueError("expression %s" % e[0])


def contraction_terms(mpo, site, left, env_keys, site_keys, blk_keys):
    """Terms (c, a, b, conj, factor) of MPO site `site`'s left (or right)
    contraction, indices into the given key lists, and the intermediate
    operators of its SumProd strings: [(dq, [(env index, conj, factor)])],
    the sum of environment operators a SumProd string multiplies
    (TensorFunctions::tensor_product, src/core/tensor_functions.hpp:
    2220-2271); a term's environment index len(env_keys) + j names
    intermediate j."""
    names = mpo["left_operator_names" if left else
                "right_operator_names"][site]["data"]
    exprs = mpo["left_operator_exprs" if left else
                "right_operator_exprs"][site]["data"]
    terms, inters = [], []
    for j, nm in enumerate(names):
        if nm[0] == "zero":
            continue
        c = blk_keys.index(key(nm))
        for x in _all_strings(exprs[j]):
            f, cj = x[1], x[2]
            if x[0] == "prod":
                a, b = x[3], x[4]
                ea = env_keys.index(key(a if left else b))
                sb = site_keys.index(key(b if left else a))
            else:
                a, b, ops, conjs = x[3], x[4], x[6], x[7]
                # the summed operators sit on the environment's side
                assert (a is None) == left and (b is None) == (not left)
                sb = site_keys.index(key(b if left else a))
                ea = len(env_keys) + len(inters)
                inters.append((ops[0][4], [(env_keys.index(key(o)), bool(cc),
                                            o[3]) for o, cc in
                                           zip(ops, conjs)]))
            terms.append((c, ea, sb, cj, f / nm[3]) if left else
                         (c, sb, ea, cj, f / nm[3]))
    return terms, inters
