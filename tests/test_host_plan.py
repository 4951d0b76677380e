"""CPU-side checks of libb2g.so: the C ABI loads, exports every symbol the
header declares, and its host plan builder reproduces the reference's
connection lists and flop count (no device is touched)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import DELAYED_FIXTURES, FIXTURES, REPO, fixture_path

import b2g
from b2g import effham
from b2g.problem import Problem


def declared_symbols():
    with open(os.path.join(REPO, "include", "b2g.h")) as f:
        txt = f.read()
    return sorted(set(re.findall(r"\b(b2g_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = b2g.lib()
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n


def test_ctypes_struct_layouts_match_the_library():
    from b2g import _lib as L
    n = len(L.ABI_STRUCTS)
    out = (C.c_size_t * n)()
    assert L.lib().b2g_abi_sizes(out, n) == n
    for name, size in zip(L.ABI_STRUCTS, out):
        assert C.sizeof(getattr(L, name)) == size, name


@pytest.mark.parametrize("name", FIXTURES)
def test_host_plan_matches_reference_connections(name):
    pb = Problem.from_npz(fixture_path(name))
    h = effham.host_plan(pb)
    try:
        g = pb.golden
        for which, pre in ((0, "cwfn"), (1, "cdiag")):
            c = effham.connections(h, which)
            for k in ("ia", "ib", "ic"):
                assert np.array_equal(c[k], g[pre + "_" + k]), (name, pre, k)
            assert np.array_equal(c["stride"], g[pre + "_stride"])
            np.testing.assert_allclose(c["factor"], g[pre + "_factor"],
                                       rtol=1e-13, atol=1e-15)
        info = effham.host_plan_info(h)
        assert info["macs"] == int(pb.meta["nflop_mac"])
        assert info["ket_len"] == len(g["psi"])
    finally:
        effham.destroy(h)


def test_host_plan_rejects_bad_operator_reference():
    pb = Problem.from_npz(fixture_path("hub8_sz"))
    pb.term_a = pb.term_a.copy()
    pb.term_a[0] = len(pb.lops) + 5
    with pytest.raises(b2g.B2GError):
        effham.host_plan(pb)


def test_host_plan_rejects_unsorted_quanta():
    pb = Problem.from_npz(fixture_path("hub8_su2"))
    q = pb.ket.quanta.copy()
    q[0], q[1] = q[1], q[0]
    pb.ket.quanta = q
    with pytest.raises(b2g.B2GError):
        effham.host_plan(pb)


def _pairs(h):
    from b2g import _lib as L
    n = C.c_uint64()
    L.check(L.lib().b2g_effham_pairs(h, C.byref(n), None))
    arr = (L.b2g_pair * max(n.value, 1))()
    L.check(L.lib().b2g_effham_pairs(h, C.byref(n), arr))
    return np.ctypeslib.as_array(arr)[:n.value]


def replay(pb, pairs, x):
    """V[window] += alpha op(A) C[window] op(B)^T for every recorded pair,
    in numpy (the pair semantics of include/b2g.h)."""
    ld, rd = pb.lops.data, pb.rops.data
    y = np.zeros(pb.bra_len)
    for q in pairs:
        cm, cn, vm, vn = (int(q[k]) for k in ("c_m", "c_n", "v_m", "v_n"))
        cld, vld = int(q["c_ld"]) or cn, int(q["v_ld"]) or vn
        co, vo, ao, bo = (int(q[k]) for k in ("c_off", "v_off", "a_off",
                                               "b_off"))
        cw = np.stack([x[co + i * cld:co + i * cld + cn] for i in range(cm)])
        a = ld[ao:ao + cm * vm]
        a = a.reshape(cm, vm).T if q["conj"] & 1 else a.reshape(vm, cm)
        b = rd[bo:bo + cn * vn]
        bt = b.reshape(cn, vn) if q["conj"] & 2 else b.reshape(vn, cn).T
        v = q["alpha"] * (a @ cw @ bt)
        for i in range(vm):
            y[vo + i * vld:vo + i * vld + vn] += v[i]
    return y


@pytest.mark.parametrize("name", DELAYED_FIXTURES)
def test_delayed_pairs_reproduce_reference_hpsi(name):
    """Three-tensor strings of a delayed operator tensor are recorded as
    rotates on row (delayed left) or column (delayed right) windows of the
    wavefunction blocks; replaying the recorded list in numpy reproduces
    the reference's own H*psi of both dumped vectors and its MAC count."""
    pb = Problem.from_npz(fixture_path(name))
    assert pb.delayed_side in (1, 2) and (pb.term_kind == 1).any()
    h = effham.host_plan(pb)
    pairs = _pairs(h)
    assert (pairs["kind"] == 1).any()
    for x, ref in ((pb.golden["psi"], pb.golden["hpsi"]),
                   (pb.golden["x2"], pb.golden["hx2"])):
        y = replay(pb, pairs, x)
        assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref), name
    assert effham.host_plan_info(h)["macs"] == int(pb.meta["nflop_mac"])
    effham.destroy(h)


@pytest.mark.parametrize("name", DELAYED_FIXTURES)
def test_delayed_connections_match_reference(name):
    pb = Problem.from_npz(fixture_path(name))
    h = effham.host_plan(pb)
    con = effham.connections(h, 0)
    g = pb.golden
    for k in ("ia", "ib", "ic", "stride"):
        assert np.array_equal(con[k], g["cwfn_" + k]), (name, k)
    assert np.allclose(con["factor"], g["cwfn_factor"], rtol=1e-14, atol=0)
    effham.destroy(h)
