"""CPU-side checks of libb2g.so: the C ABI loads, exports every symbol the
header declares, and its host plan builder reproduces the reference's
connection lists and flop count (no device is touched)."""
import os
import re

import numpy as np
import pytest

from conftest import FIXTURES, REPO, fixture_path

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
