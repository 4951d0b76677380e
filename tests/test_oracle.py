"""Pins the oracle (oracle/b2g_oracle.c) before anything is checked against it.

* golden vectors dumped by the reference itself (oracle/_ref harness, see
  tests/golden/make_fixtures.py): ConnectionInfo lists, H*psi of two vectors,
  the diagonal, the Davidson energy, the reference's flop count;
* brute force on a tiny site: dense H from unit vectors -> symmetry,
  diagonal, lowest eigenvalue;
* the SU(2) symbols against definitions that do not share the Racah 6j
  formula (3j-symbol sums) and textbook identities.
"""
import itertools
import json
import math
from fractions import Fraction

import numpy as np
import pytest

from conftest import FIXTURES, fixture_path

from b2g.problem import Problem
import b2g_oracle as O


def load(name):
    return Problem.from_npz(fixture_path(name))


@pytest.fixture(scope="module", params=FIXTURES)
def case(request):
    pb = load(request.param)
    return request.param, pb, O.Oracle(pb)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_connections_match_reference(case):
    name, pb, orc = case
    g = pb.golden
    for which, pre in ((0, "cwfn"), (1, "cdiag")):
        c = orc.connections(which)
        assert np.array_equal(c["n"], g[pre + "_n"]), name
        assert np.array_equal(c["quanta"], g[pre + "_quanta"])
        assert np.array_equal(c["idx"], g[pre + "_idx"])
        for k in ("ia", "ib", "ic"):
            assert np.array_equal(c[k], g[pre + "_" + k]), (name, pre, k)
        assert np.array_equal(c["stride"], g[pre + "_stride"])
        np.testing.assert_allclose(c["factor"], g[pre + "_factor"], rtol=1e-13,
                                   atol=1e-15)


def test_hpsi_matches_reference(case):
    name, pb, orc = case
    g = pb.golden
    assert rel(orc.hpsi(g["psi"]), g["hpsi"]) < 1e-12
    assert rel(orc.hpsi(g["x2"], nthreads=4), g["hx2"]) < 1e-12


def test_flop_count_matches_reference(case):
    name, pb, orc = case
    assert orc.macs == int(pb.meta["nflop_mac"])


def test_diag_matches_reference(case):
    name, pb, orc = case
    d = orc.diag()
    assert rel(d, pb.golden["diag"]) < 1e-13


def test_davidson_matches_reference(case):
    name, pb, orc = case
    e, x, nd, rc = orc.davidson(pb.golden["psi"], conv=1e-10)
    assert rc == 0
    assert abs(e[0] - float(pb.meta["dav_e"])) < 1e-8
    # same start vector, same steps: the reference's iteration count
    assert nd == int(pb.meta["dav_ndav"]), (name, nd)
    # the Ritz vector is an eigenvector
    hx = orc.hpsi(x)
    assert np.linalg.norm(hx - e[0] * x) < 1e-4


def test_bruteforce_dense_hamiltonian():
    pb = load("hub8_su2")
    orc = O.Oracle(pb)
    n = orc.ket_len
    assert n < 2000
    H = np.empty((n, n))
    e = np.zeros(n)
    for i in range(n):
        e[i] = 1.0
        H[:, i] = orc.hpsi(e)
        e[i] = 0.0
    assert np.abs(H - H.T).max() < 1e-12
    assert np.abs(np.diag(H) - orc.diag()).max() < 1e-13
    w = np.linalg.eigvalsh(H)
    assert abs(w[0] - float(pb.meta["dav_e"])) < 1e-9
    # H*psi of the fixture vector through the dense matrix
    assert rel(H @ pb.golden["psi"], pb.golden["hpsi"]) < 1e-12


@pytest.fixture(scope="module")
def dense_hub8():
    """Dense H_eff of the Hubbard L=8 SU(2) site (unit vectors through the
    oracle) and its full spectrum (numpy eigh)."""
    pb = load("hub8_su2")
    orc = O.Oracle(pb)
    n = orc.ket_len
    H = np.empty((n, n))
    e = np.zeros(n)
    for i in range(n):
        e[i] = 1.0
        H[:, i] = orc.hpsi(e)
        e[i] = 0.0
    w, v = np.linalg.eigh(H)
    return pb, orc, H, w, v


def _guess(pb, k, seed=3):
    rng = np.random.default_rng(seed)
    x = np.array([pb.golden["psi"]] + [rng.standard_normal(len(
        pb.golden["psi"])) for _ in range(k - 1)])
    return x


def test_davidson_k_roots_against_eigh(dense_hub8):
    """k = 4 roots: the lowest four eigenvalues of the dense matrix, Ritz
    vectors orthonormal eigenvectors."""
    pb, orc, H, w, v = dense_hub8
    e, x, nd, rc = orc.davidson(_guess(pb, 4), conv=1e-11)
    assert rc == 0
    assert np.abs(e - w[:4]).max() < 1e-9, (e, w[:4])
    assert np.abs(x @ x.T - np.eye(4)).max() < 1e-8
    for i in range(4):
        assert np.linalg.norm(H @ x[i] - e[i] * x[i]) < 1e-4


def test_davidson_dependent_guesses(dense_hub8):
    """Two identical guesses: the subspace starts from one vector (m < k)
    and roots may only converge once m >= k; both roots still come out."""
    pb, orc, H, w, v = dense_hub8
    x0 = np.array([pb.golden["psi"], pb.golden["psi"]])
    e, x, nd, rc = orc.davidson(x0, conv=1e-11)
    assert rc == 0
    assert np.abs(e - w[:2]).max() < 1e-9, (e, w[:2])


def test_davidson_orthogonal_states(dense_hub8):
    """ors = the ground state: projected out, the solver finds the first
    excited state; with projection weight 10 (H + 10|v0><v0|) likewise."""
    pb, orc, H, w, v = dense_hub8
    g = v[:, 0]
    e, x, nd, rc = orc.davidson(pb.golden["psi"], conv=1e-11, ors=[g])
    assert rc == 0 and abs(e[0] - w[1]) < 1e-9, (e, w[:3])
    assert abs(float(x @ g)) < 1e-8
    e, x, nd, rc = orc.davidson(pb.golden["psi"], conv=1e-11, ors=[g],
                                proj_weights=[10.0])
    assert rc == 0 and abs(e[0] - w[1]) < 1e-9, (e, w[:3])


@pytest.mark.parametrize("dtype,pick", [
    (4, lambda w, s: w[np.argmin(np.abs(w - s))]),        # CloseTo
    (2, lambda w, s: w[w <= s].max()),                     # LessThan
    (1, lambda w, s: w[w > s].min()),                      # GreaterThan
])
def test_davidson_shift_types(dense_hub8, dtype, pick):
    """CloseTo / LessThan / GreaterThan a shift between the third and the
    fourth eigenvalue converge to the eigenvalue the type selects, from a
    start vector that mixes the third and fourth eigenvectors with noise.
    (LessThan ranks Ritz values at or below the shift first, closest first,
    then those above it, farthest first: from a start with no component
    below the shift it climbs to the top of the spectrum, as the reference
    does.)"""
    pb, orc, H, w, v = dense_hub8
    shift = 0.3 * w[2] + 0.7 * w[3]
    x0 = (v[:, 2] + v[:, 3]) / np.sqrt(2) + 0.05 * np.random.default_rng(
        11).standard_normal(len(w)) / np.sqrt(len(w))
    e, x, nd, rc = orc.davidson(x0, conv=1e-11,
                                davidson_type=dtype, shift=shift,
                                max_iter=2000)
    assert rc == 0
    assert abs(e[0] - pick(w, shift)) < 1e-8, (dtype, e, w[:6])


@pytest.mark.parametrize("dtype", [32, 64])
def test_davidson_preconditioners(dense_hub8, dtype):
    """Davidson's diagonal preconditioner and no preconditioner reach the
    same ground state as Olsen's."""
    pb, orc, H, w, v = dense_hub8
    e, x, nd, rc = orc.davidson(pb.golden["psi"], conv=1e-11,
                                davidson_type=dtype, max_iter=2000)
    assert rc == 0 and abs(e[0] - w[0]) < 1e-9


def test_davidson_soft_max_iter_and_rel_conv(dense_hub8):
    pb, orc, H, w, v = dense_hub8
    e, x, nd, rc = orc.davidson(pb.golden["psi"], conv=1e-14,
                                soft_max_iter=3)
    assert rc == 2 and nd == 3
    # a relative threshold of 1e-2 * |E| on the residual norm stops earlier
    e0, _, nd0, _ = orc.davidson(pb.golden["psi"], conv=1e-12)
    e1, _, nd1, rc1 = orc.davidson(pb.golden["psi"], conv=0.0, rel_conv=1e-2)
    assert rc1 == 0 and nd1 < nd0 and abs(e1[0] - w[0]) < 1e-3


def test_planted_mistake_is_caught():
    """Dropping one string of H_eff must move H*psi far beyond tolerance."""
    pb = load("n2_su2")
    ref = pb.golden["hpsi"]
    pb.term_factor = pb.term_factor.copy()
    pb.term_factor[len(pb.term_factor) // 2] = 0.0
    y = O.Oracle(pb).hpsi(pb.golden["psi"])
    assert rel(y, ref) > 1e-6


@pytest.mark.parametrize("name", ["n2_su2", "hub8_sz"])
def test_grad_ops_linearity_and_euler(name):
    """E(L, R) = <v|H(L, R)|c> is linear in the left-operator data for fixed
    right data and vice versa, so the oracle's partial expectations must give
    E(L + dL, R) - E(L, R) = <g_L, dL> exactly (no finite-difference error),
    the same for R, and Euler's identity E = <g_L, L> = <g_R, R>.  A sign,
    index or transpose slip in either gradient fails these."""
    import copy
    pb = load(name)
    rng = np.random.default_rng(17)
    orc = O.Oracle(pb)
    v = rng.standard_normal(orc.bra_len)
    c = rng.standard_normal(orc.ket_len)
    e0 = float(v @ orc.hpsi(c))
    gl, gr = orc.grad_ops(v, c)
    tol = 1e-10 * max(1.0, abs(e0))
    assert abs(float(gl @ pb.lops.data) - e0) < tol
    assert abs(float(gr @ pb.rops.data) - e0) < tol
    for side in ("l", "r"):
        pb2 = copy.copy(pb)
        ops = copy.copy(pb.lops if side == "l" else pb.rops)
        d = rng.standard_normal(len(ops.data))
        ops.data = ops.data + d
        if side == "l":
            pb2.lops = ops
        else:
            pb2.rops = ops
        e1 = float(v @ O.Oracle(pb2).hpsi(c))
        g = gl if side == "l" else gr
        assert abs((e1 - e0) - float(g @ d)) < 1e-10 * max(1.0, abs(e1)), side


def test_reference_energies_known_answers():
    with open(fixture_path("energies").replace(".npz", ".json")) as f:
        en = json.load(f)
    # docs/source/tutorial/hubbard.ipynb: L=8, U=2 half filling
    assert abs(en["hub8_sz"]["final"] - -6.225634144662398) < 1e-8
    assert abs(en["hub8_su2"]["final"] - -6.225634144676156) < 1e-8
    # pyblock2/main_test/000-check.py, unit_test/test_dmrg_n2_sto3g.cpp:191
    assert abs(en["n2_su2_m500"]["final"] - -107.654122447525) < 1e-8
    # SZ and SU(2) of the same Hamiltonian agree
    assert abs(en["hub8_sz"]["final"] - en["hub8_su2"]["final"]) < 1e-9


# ----------------------------------------------------------- SU(2) symbols
def _f(n):
    return math.factorial(n)


def threej(j1, j2, j3, m1, m2, m3):
    """Wigner 3j symbol of doubled spins from Racah's single-sum formula
    (exact rational arithmetic, square root at the end)."""
    if m1 + m2 + m3 != 0:
        return 0.0
    if (j1 + j2 + j3) % 2 or j3 > j1 + j2 or j3 < abs(j1 - j2):
        return 0.0
    if any(abs(m) > j for m, j in ((m1, j1), (m2, j2), (m3, j3))):
        return 0.0
    if any((j - m) % 2 for m, j in ((m1, j1), (m2, j2), (m3, j3))):
        return 0.0
    J1, J2, J3 = Fraction(j1, 2), Fraction(j2, 2), Fraction(j3, 2)
    M1, M2, M3 = Fraction(m1, 2), Fraction(m2, 2), Fraction(m3, 2)
    tri = Fraction(_f(int(J1 + J2 - J3)) * _f(int(J1 - J2 + J3)) *
                   _f(int(-J1 + J2 + J3)), _f(int(J1 + J2 + J3 + 1)))
    pre = tri * (_f(int(J1 + M1)) * _f(int(J1 - M1)) * _f(int(J2 + M2)) *
                 _f(int(J2 - M2)) * _f(int(J3 + M3)) * _f(int(J3 - M3)))
    s = Fraction(0)
    for k in range(0, int(J1 + J2 + J3) + 2):
        d = [k, J1 + J2 - J3 - k, J1 - M1 - k, J2 + M2 - k, J3 - J2 + M1 + k,
             J3 - J1 - M2 + k]
        if any(x < 0 for x in d):
            continue
        den = 1
        for x in d:
            den *= _f(int(x))
        s += Fraction((-1) ** k, den)
    sign = (-1) ** int(J1 - J2 - M3)
    return float(sign * s) * math.sqrt(float(pre))


def sixj_brute(a, b, c, d, e, f):
    """6j symbol as the sum over all projections of four 3j symbols,
    {j1 j2 j3; j4 j5 j6} = sum (-1)^S (j1 j2 j3; -m1 -m2 -m3)
    (j1 j5 j6; m1 -m5 m6) (j4 j2 j6; m4 m2 -m6) (j4 j5 j3; -m4 m5 m3),
    S = sum_k (j_k - m_k) (doubled spins throughout)."""
    r = 0.0
    rng = [range(-j, j + 1, 2) for j in (a, b, c, d, e, f)]
    for m1, m2, m3 in itertools.product(rng[0], rng[1], rng[2]):
        x1 = threej(a, b, c, -m1, -m2, -m3)
        if x1 == 0.0:
            continue
        for m4, m5, m6 in itertools.product(rng[3], rng[4], rng[5]):
            x2 = threej(a, e, f, m1, -m5, m6)
            if x2 == 0.0:
                continue
            x3 = threej(d, b, f, m4, m2, -m6)
            x4 = threej(d, e, c, -m4, m5, m3)
            s = (a - m1 + b - m2 + c - m3 + d - m4 + e - m5 + f - m6) // 2
            r += (-1) ** s * x1 * x2 * x3 * x4
    return r


@pytest.mark.parametrize("args", [(1, 1, 2, 1, 1, 0), (2, 2, 2, 2, 2, 2),
                                  (1, 2, 1, 2, 1, 2), (3, 1, 2, 1, 3, 2),
                                  (2, 4, 2, 3, 1, 3), (4, 2, 2, 2, 2, 4)])
def test_sixj_bruteforce(args):
    assert abs(O.sixj(*args) - sixj_brute(*args)) < 1e-13


def test_sixj_identities():
    rng = np.random.default_rng(7)
    for _ in range(200):
        a, b = (int(x) for x in rng.integers(0, 20, 2))
        f = abs(a - b) + 2 * int(rng.integers(0, min(a, b) + 1))
        # sum_x (-1)^(2x) (2x+1) {a b x; a b f} = 1   (Messiah C.35a)
        s = sum(((-1) ** x) * (x + 1) * O.sixj(a, b, x, a, b, f)
                for x in range(abs(a - b), a + b + 1))
        assert abs(s - 1) < 1e-12
    # {j j' 0; J J' g} closed form
    for _ in range(200):
        j, J = rng.integers(0, 20, 2)
        g = int(rng.integers(abs(j - J), j + J + 1))
        if (j + J + g) % 2:
            continue
        v = O.sixj(int(j), int(j), 0, int(J), int(J), g)
        exp = (-1) ** ((j + J + g) // 2) / math.sqrt((j + 1) * (J + 1))
        assert abs(v - exp) < 1e-13


def test_ninej_zero_identity():
    """{a b e; c d e; f f 0} = (-1)^(b+c+e+f) {a b e; d c f} / sqrt((2e+1)(2f+1))
    (Edmonds 6.4.14)."""
    rng = np.random.default_rng(3)
    n = 0
    while n < 200:
        a, b, c, d = (int(x) for x in rng.integers(0, 12, 4))
        e = int(rng.integers(0, 12))
        f = int(rng.integers(0, 12))
        if (a + b + e) % 2 or (c + d + e) % 2 or (a + c + f) % 2 or \
                (b + d + f) % 2:
            continue
        v = O.ninej(a, b, e, c, d, e, f, f, 0)
        exp = (-1) ** ((b + c + e + f) // 2) * O.sixj(a, b, e, d, c, f) / \
            math.sqrt((e + 1) * (f + 1))
        assert abs(v - exp) < 1e-13
        n += 1
