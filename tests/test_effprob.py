"""block2's EFFPROB1 effective-problem files (b2g.effprob).

The harness writes each dumped site's effective Hamiltonian with the
reference's own EffectiveProblem::save_problem_file (what
EffectiveHamiltonian::save_eff_problem calls); the fixtures keep those
bytes next to the arrays the harness dumped from the same objects.
Checked on the CPU:
  * the Problem read from the file equals the dumped one: ket / bra and
    left / right operator infos, the H strings (conjugation, factor), for
    every string the quantum, block structure and data presence of the
    operators (and delayed sub-operators) it uses, and the tensor-product
    connection lists; the dumped operator factors are 1 (the file does not
    hold them).  The dump numbers operators by object, the reader by value
    (each operator once, by first use), so operators are compared through
    the strings that use them;
  * the wavefunction ConnectionInfo stored in the file equals the dumped
    one;
  * with the dumped operator data attached (matched through the strings),
    the host plan of the file's
    problem replayed in numpy reproduces the reference's H*psi of both
    dumped vectors (1e-12 relative) with the reference's MAC count;
  * the environments' OperatorTensor files (OperatorTensor::save_data)
    supply exactly the dumped operator numbers and factors, and the
    problem assembled from the three files reproduces the reference's
    H*psi;
  * damaged files (bad magic, truncated, trailing bytes) and missing
    operator data are refused.
"""
import numpy as np
import pytest

from conftest import DELAYED_FIXTURES, FIXTURES, fixture_path

from b2g import effham
from b2g import effprob as EP
from b2g.problem import Problem
from test_host_plan import _pairs, replay

ALL = FIXTURES + DELAYED_FIXTURES


def load(name):
    pb = Problem.from_npz(fixture_path(name))
    if "eff_problem_smat" not in pb.golden:
        pytest.skip("fixture without an EFFPROB1 file")
    return pb, bytes(pb.golden["eff_problem_smat"])


def same_infos(a, b):
    for k in ("dq", "fermion", "n", "off", "quanta", "nbra", "nket", "shift"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def info_blocks(iset, i):
    o, n = int(iset.off[i]), int(iset.n[i])
    return (int(iset.dq[i]), int(iset.fermion[i]),
            iset.quanta[o:o + n].tolist(), iset.nbra[o:o + n].tolist(),
            iset.nket[o:o + n].tolist(), iset.shift[o:o + n].tolist())


def op_sides(p):
    """(ops, infos, own, term index array) of the four operator lists"""
    sides = [(p.lops, p.linfo, False, p.term_a),
             (p.rops, p.rinfo, False, p.term_b)]
    if p.delayed_side:
        sides += [(p.dlops, p.dlinfo, True, p.term_da),
                  (p.drops, p.drinfo, True, p.term_db)]
    return sides


def op_len(ops, infos, own, k):
    if ops.flags[k] & 1:
        return 0
    return infos.total(k if own else int(ops.info[k]))


def attach(fp, pb):
    """Copy the dumped operator data into the file problem's buffers,
    matching operators through the H strings that use them."""
    bufs = [np.zeros(fp.data_len[0]), np.zeros(fp.data_len[1])]
    src = [pb.lops.data, pb.rops.data]
    ds = fp.delayed_side - 1
    for si, ((fo, fi, own, ft), (po, pi, _, pt)) in enumerate(
            zip(op_sides(fp), op_sides(pb))):
        side = si if si < 2 else ds
        for t in range(len(ft)):
            k, j = int(ft[t]), int(pt[t])
            if k < 0:
                continue
            n = op_len(fo, fi, own, k)
            a, b = int(fo.off[k]), int(po.off[j])
            bufs[side][a:a + n] = src[side][b:b + n]
    fp.lops.data, fp.rops.data = bufs


@pytest.mark.parametrize("name", ALL)
def test_effprob_structure_matches_dump(name):
    pb, buf = load(name)
    f = EP.EffProblemFile(buf)
    assert f.sym == pb.sym and f.opdq == pb.opdq
    fp = EP.to_problem(f)
    same_infos(fp.ket, pb.ket)
    same_infos(fp.bra, pb.bra)
    same_infos(fp.linfo, pb.linfo)
    same_infos(fp.rinfo, pb.rinfo)
    for k in ("term_conj", "term_factor", "term_kind", "term_dconj"):
        assert np.array_equal(getattr(fp, k), getattr(pb, k)), k
    assert fp.delayed_side == pb.delayed_side
    # the operators each string uses: quantum, structure, data or not
    for (fo, fi, own, ft), (po, pi, _, pt) in zip(op_sides(fp), op_sides(pb)):
        assert (po.factor == 1).all()
        assert np.array_equal(ft < 0, pt < 0)
        for t in range(len(ft)):
            k, j = int(ft[t]), int(pt[t])
            if k < 0:
                continue
            assert fo.dq[k] == po.dq[j], t
            assert fo.flags[k] == po.flags[j], t
            assert (info_blocks(fi, k if own else int(fo.info[k])) ==
                    info_blocks(pi, j if own else int(po.info[j]))), t
        # numbered by first use, each operator once
        used = [int(x) for x in ft if x >= 0]
        firsts = list(dict.fromkeys(used))
        assert firsts == list(range(len(fo.dq)))
    if pb.delayed_side:
        assert len(fp.tp) == len(pb.tp)
        for a, b in zip(fp.tp, pb.tp):
            assert np.array_equal(a["n"], b["n"]) and a["nc"] == b["nc"]
            for k in ("quanta", "idx", "ia", "ib", "ic", "stride", "factor"):
                assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("name", ALL)
def test_effprob_wavefunction_connections_match_dump(name):
    pb, buf = load(name)
    f = EP.EffProblemFile(buf)
    i = f.operator_quanta.index(f.opdq)
    c = f.wfn_cinfos[i]
    g = pb.golden
    for k in ("ia", "ib", "ic", "stride", "factor"):
        assert np.array_equal(c[k], g["cwfn_" + k]), k


@pytest.mark.parametrize("name", ALL)
def test_effprob_problem_reproduces_reference_hpsi(name):
    pb, buf = load(name)
    fp = EP.to_problem(EP.EffProblemFile(buf))
    attach(fp, pb)
    h = effham.host_plan(fp)
    try:
        pairs = _pairs(h)
        for x, ref in ((pb.golden["psi"], pb.golden["hpsi"]),
                       (pb.golden["x2"], pb.golden["hx2"])):
            y = replay(fp, pairs, x)
            assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)
        assert effham.host_plan_info(h)["macs"] == int(pb.meta["nflop_mac"])
    finally:
        effham.destroy(h)


def tensors(pb):
    g = pb.golden
    return [EP.tensor_from_bytes(bytes(g[k])) if k in g else None
            for k in ("lopt_smat", "ropt_smat", "dlopt_smat", "dropt_smat")]


@pytest.mark.parametrize("name", ALL)
def test_operator_tensor_files_complete_the_problem(name):
    """The environments' OperatorTensor files (OperatorTensor::save_data,
    written by the harness next to the EFFPROB1 file) fill the file
    problem's operator buffers with exactly the dumped numbers and factors,
    and the assembled problem reproduces the reference's H*psi."""
    pb, buf = load(name)
    if "lopt_smat" not in pb.golden:
        pytest.skip("fixture without operator-tensor files")
    fp = EP.to_problem(EP.EffProblemFile(buf))
    EP.attach_operator_data(fp, *tensors(pb))
    ref = EP.to_problem(EP.EffProblemFile(buf))
    attach(ref, pb)
    assert np.array_equal(fp.lops.data, ref.lops.data)
    assert np.array_equal(fp.rops.data, ref.rops.data)
    for (fo, _, _, ft), (po, _, _, pt) in zip(op_sides(fp), op_sides(pb)):
        for t in range(len(ft)):
            if ft[t] >= 0:
                assert fo.factor[ft[t]] == po.factor[pt[t]]
    h = effham.host_plan(fp)
    try:
        y = replay(fp, _pairs(h), pb.golden["psi"])
        ref_y = pb.golden["hpsi"]
        assert np.linalg.norm(y - ref_y) <= 1e-12 * np.linalg.norm(ref_y)
    finally:
        effham.destroy(h)


def test_operator_tensor_attach_refuses_missing_data():
    pb, buf = load("hub8_sz")
    fp = EP.to_problem(EP.EffProblemFile(buf))
    left, right = tensors(pb)[:2]
    with pytest.raises(EP.FormatError, match="has no data"):
        EP.attach_operator_data(fp, left, dict(right, ops=right["ops"][1:]))
    with pytest.raises(EP.FormatError, match="after the operator tensor"):
        EP.tensor_from_bytes(bytes(pb.golden["lopt_smat"]) + b"\0")


def test_effprob_refuses_damaged_files():
    _, buf = load("hub8_sz")
    with pytest.raises(EP.FormatError, match="not an EFFPROB1"):
        EP.EffProblemFile(b"EFFPROB2" + buf[8:])
    with pytest.raises(EP.FormatError, match="ends inside"):
        EP.EffProblemFile(buf[:len(buf) // 2])
    with pytest.raises(EP.FormatError, match="after the EFFPROB1"):
        EP.EffProblemFile(buf + b"\0")
