"""The truncation oracle (oracle/truncate_oracle.py) against the reference.

The golden fixtures hold, for the Davidson vector of each dumped site, the
reference's own MovingEnvironment::density_matrix and split_density_matrix
results (oracle/ref_harness.cpp dump_split) in five cases: trace_right
Physical (sR), trace_left Physical (sL), trace_left Reduced | KeepOne (sK),
and with noise 1e-6 drawn from block2's generator seeded with 4321:
wavefunction noise, trace_right (sW), density-matrix noise, trace_left
(sD).  The sweep fixtures add, at the sites dumped along block2's sweeps,
density-matrix noise with every state kept, trace_right (sE) and
trace_left (sF).  The noisy density matrices pin the oracle's restatement
of block2's random stream (std::mt19937 through
uniform_real_distribution) and of where the noise enters.
Checked here, on the CPU:
  * density-matrix structure exact and data to 1e-13 (absolute; the
    density matrix of a normalised vector has trace 1);
  * truncation error to 1e-12 and the squared spectra to 1e-13 (absolute;
    sqrt of eigenvalues near zero is not compared directly);
  * the factors' block structure exactly, and sign-free quantities of the
    factors: the truncated wavefunction left * right, block by block, to
    1e-8 (the rotation basis inside a degenerate kept space is open; kept
    states whose eigenvalues lie at the rounding level of ||D||, eps ||D||
    ~ 1e-16 .. 1e-14, are fixed only up to rotations among themselves by
    any FP64 eigensolver and carry weight sqrt(lambda) ~ 1e-8 in the
    wavefunction: 1e-9-level differences between numpy's LAPACK and the
    reference's MKL on the L=16 Hubbard site);
  * independent of the reference: error + kept weight = trace of the
    density matrix, orthonormal rotation matrices, eigen residuals of the
    oracle's eigenpairs, and a planted mistake (a sign error in the
    selection) that the reference comparison catches.
"""
import json

import numpy as np
import pytest

import truncate_oracle as T
from conftest import FIXTURES, DELAYED_FIXTURES, fixture_path
from sweep_common import (SWEEP_SPLIT_CASES, load_sweep_split,
                          sweep_split_fixtures)

CASES = ["sR", "sL", "sK", "sW", "sD"]


def settings(z, tag):
    """err, k, cutoff, trunc_type, trace_right, noise, noise_type, seed"""
    st = [float(x) for x in z[tag + "_settings"]]
    return st + [0.0, 0.0, 0.0][:8 - len(st)]
ALL = FIXTURES + DELAYED_FIXTURES


def load(name):
    z = np.load(fixture_path(name))
    meta = json.loads(bytes(z["meta_json"]).decode())
    sym = T.SU2 if meta["sym"] == "su2" else T.SZ
    info = dict(quanta=z["ket_quanta"], nbra=z["ket_nbra"],
                nket=z["ket_nket"], shift=z["ket_shift"],
                dq=int(z["ket_dq"][0]))
    return z, sym, info


def ref_part(z, tag, part):
    p = tag + "_" + part
    return (z[p + "_quanta"].astype(np.uint64), z[p + "_nbra"],
            z[p + "_nket"], z[p + "_shift"], z[p + "_data"])


def blocks_of(quanta, nbra, nket, shift, data):
    return {int(q): data[int(o):int(o) + int(m) * int(n)].reshape(int(m), int(n))
            for q, m, n, o in zip(quanta, nbra, nket, shift)}


def recon(sym, info, res_left, res_right, trace_right):
    """dense reconstruction of the truncated two-site wavefunction:
    dict wavefunction quantum -> (n_bra x n_ket) block"""
    qs = [int(q) for q in info["quanta"]]
    dq = info["dq"]
    lab = {q: (T.q_get_bra(sym, q, dq) if trace_right
               else T.q_neg(sym, T.q_get_ket(sym, q))) for q in qs}
    rot, wav = (res_left, res_right) if trace_right else (res_right, res_left)
    out = {}
    for q, w in wav.items():
        r = rot[lab[q]]
        out[q] = r @ w if trace_right else w @ r
    return out


def oracle_parts(res):
    def as_dict(p):
        q, nb, nk, data = p
        return {int(a): np.asarray(d).reshape(m, n)
                for a, m, n, d in zip(q, nb, nk, data)}
    return as_dict(res["left"]), as_dict(res["right"])


@pytest.mark.parametrize("name", ALL)
@pytest.mark.parametrize("tag", CASES)
def test_oracle_split_matches_reference(name, tag):
    z, sym, info = load(name)
    if tag + "_settings" not in z.files:
        pytest.skip("fixture without truncation dump")
    check_oracle_split(z, sym, info, name, tag)


@pytest.mark.parametrize("case,env", sweep_split_fixtures())
@pytest.mark.parametrize("tag", SWEEP_SPLIT_CASES)
def test_oracle_split_matches_sweep_dumps(case, env, tag):
    """The noisy truncations dumped along block2's sweeps (every state
    kept, sE / sF: the states dropped are those the noise leaves below the
    cutoff)."""
    z, sym, info = load_sweep_split(case, env)
    check_oracle_split(z, sym, info, case + "/" + env, tag)


def check_oracle_split(z, sym, info, name, tag):
    err_ref, k, cutoff, tt, tr, noise, nt, seed = settings(z, tag)
    psi = z["psi_dav"]
    res = T.split_density_matrix(sym, info, psi, int(k), cutoff, bool(tr),
                                 True, int(tt), 1.0, noise, int(nt),
                                 int(seed))
    # density matrix
    dq, dn, dnk, dsh, ddata = ref_part(z, tag, "dm")
    uq, dims, blocks = res["dm"]
    assert [int(x) for x in dq] == uq and list(dn) == dims
    for i, b in enumerate(blocks):
        o = int(dsh[i])
        ref = ddata[o:o + dims[i] ** 2].reshape(dims[i], dims[i])
        assert np.abs(b - ref).max(initial=0) < 1e-13, (name, tag, i)
    # error and spectra
    assert abs(res["error"] - err_ref) < 1e-12, (name, tag, res["error"], err_ref)
    sp = np.array(res["spectra"])
    assert sp.shape == z[tag + "_spectra"].shape
    assert np.abs(sp ** 2 - z[tag + "_spectra"] ** 2).max(initial=0) < 1e-13
    # factor structure
    for part, mine in (("left", res["left"]), ("right", res["right"])):
        q, nb, nk, _, _ = ref_part(z, tag, part)
        assert [int(x) for x in q] == list(mine[0]), (name, tag, part)
        assert list(nb) == list(mine[1]) and list(nk) == list(mine[2])
    # truncated wavefunction (sign-free)
    lref = blocks_of(*ref_part(z, tag, "left"))
    rref = blocks_of(*ref_part(z, tag, "right"))
    lo, ro = oracle_parts(res)
    a = recon(sym, info, lref, rref, bool(tr))
    b = recon(sym, info, lo, ro, bool(tr))
    assert a.keys() == b.keys()
    for q in a:
        assert np.abs(a[q] - b[q]).max(initial=0) < 1e-8, (name, tag, q)


@pytest.mark.parametrize("name", ["hub8_su2", "n2_sz", "hub16_sz_m200"])
def test_oracle_split_invariants(name):
    z, sym, info = load(name)
    psi = z["psi_dav"]
    for tr in (True, False):
        k_all = sum(int(m if tr else n) for m, n in zip(info["nbra"], info["nket"]))
        res = T.split_density_matrix(sym, info, psi, max(1, k_all // 3),
                                     1e-14, tr, False, 0)
        trace = sum(float(np.trace(b)) for b in res["dm"][2])
        kept = sum(float(res["eigvals"][b][j]) for b, j in res["ss"])
        assert abs(res["error"] + kept - trace) < 1e-12
        # eigenpairs of every block
        for blk in res["dm"][2]:
            w, v = T.signed_eigh(blk)
            assert np.abs(blk @ v.T - v.T * w).max(initial=0) < 1e-12
            assert np.abs(v @ v.T - np.eye(len(w))).max(initial=0) < 1e-12
        # rotation matrices have orthonormal kept vectors
        rot = res["left"] if tr else res["right"]
        for d in rot[3]:
            g = d.T @ d if tr else d @ d.T
            assert np.abs(g - np.eye(g.shape[0])).max(initial=0) < 1e-12
        # without normalisation, the kept weight is the truncated norm
        lo, ro = oracle_parts(res)
        rec = recon(sym, info, lo, ro, tr)
        assert abs(sum(float((x * x).sum()) for x in rec.values()) - kept) < 1e-12


def test_planted_selection_mistake_is_caught():
    """A selection that keeps the smallest instead of the largest states
    (a dropped minus sign in the sort) must fail the reference check."""
    z, sym, info = load("hub8_su2")
    err_ref, k, cutoff, tt, tr = z["sR_settings"]
    orig = T.truncate

    def wrong(sym_, quanta, w, k_, cutoff_, tt_):
        neg = [-x for x in w]
        ss, _ = orig(sym_, quanta, neg, k_, -1e300, tt_)
        err = sum(float(w[b][j]) for b in range(len(w))
                  for j in range(len(w[b])) if (b, j) not in set(ss))
        return ss, err
    T.truncate = wrong
    try:
        res = T.split_density_matrix(sym, info, z["psi_dav"], int(k), cutoff,
                                     bool(tr), True, int(tt))
    finally:
        T.truncate = orig
    assert abs(res["error"] - err_ref) > 1e-6


def test_noise_depends_on_the_seed():
    """a different seed gives a different noisy density matrix: the
    reference comparison of the noisy cases checks the random stream"""
    z, sym, info = load("hub8_sz")
    err_ref, k, cutoff, tt, tr, noise, nt, seed = settings(z, "sD")
    a = T.split_density_matrix(sym, info, z["psi_dav"], int(k), cutoff,
                               bool(tr), True, int(tt), 1.0, noise, int(nt),
                               int(seed) + 1)
    sh, dd = z["sD_dm_shift"], z["sD_dm_data"]
    uq, dims, blocks = a["dm"]
    diff = max(np.abs(b - dd[int(sh[i]):int(sh[i]) + dims[i] ** 2].reshape(
        dims[i], dims[i])).max(initial=0) for i, b in enumerate(blocks))
    assert diff > 1e-10


def test_label_algebra():
    """get_bra / get_ket / negation against values the reference's label
    layout fixes (SU2LongLong: N bits 48-63, 2S_low 32-47, 2S 16-31, pg
    0-15)."""
    def su2(n, lo, tw, pg=0):
        return ((n & 0xFFFF) << 48) | (lo << 32) | (tw << 16) | pg
    q = su2(4, 1, 3, 2)
    assert T.q_get_ket(T.SU2, q) == su2(4, 3, 3, 2)
    assert T.q_get_bra(T.SU2, q, su2(1, 1, 1, 3)) == su2(5, 1, 1, 1)
    assert T.q_neg(T.SU2, q) == su2(-4, 1, 3, 2)
    assert T.multiplicity(T.SU2, q) == 4

    def sz(n, tz, pg=0):
        return ((n & 0xFFFF) << 48) | ((tz & 0xFFFF) << 16) | pg
    assert T.q_neg(T.SZ, sz(3, -1, 1)) == sz(-3, 1, 1)
    assert T.q_get_bra(T.SZ, sz(3, -1), sz(1, 1)) == sz(4, 0)
    assert T.q_get_ket(T.SZ, sz(3, 1)) == sz(3, 1)


def wav_blocks(info_part, data):
    q, nb, nk, sh = info_part
    return {int(a): data[int(o):int(o) + int(m) * int(n)].reshape(int(m), int(n))
            for a, m, n, o in zip(q, nb, nk, sh)}


@pytest.mark.parametrize("name", ALL)
@pytest.mark.parametrize("tag", ["sM", "sN", "sQ"])
def test_oracle_multi_split_matches_reference(name, tag):
    """State-averaged split of two roots (the Davidson vector and a
    normalised random vector of the same structure, weights 0.7 / 0.3):
    MovingEnvironment::density_matrix_with_multi_target +
    multi_split_density_matrix (moving_environment.hpp:3582-3608,
    4347-4462), trace_right, Physical, k = 2/3 of the states; without
    noise (sM), with wavefunction (sN) and density-matrix (sQ) noise 1e-6
    from block2's generator seeded with 4321 (one tensor per root from the
    continuing stream)."""
    z, sym, info = load(name)
    if tag + "_settings" not in z.files:
        pytest.skip("fixture without multi-root dump")
    err_ref, k, cutoff, tt, tr, noise, nt, seed = settings(z, tag)
    psis = [z["psi_dav"], z[tag + "_roots"]]
    res = T.multi_split_density_matrix(sym, info, psis, [0.7, 0.3], int(k),
                                       cutoff, bool(tr), True, int(tt), 1.0,
                                       noise, int(nt), int(seed))
    dq, dn, dnk, dsh, ddata = ref_part(z, tag, "dm")
    uq, dims, blocks = res["dm"]
    assert [int(x) for x in dq] == uq and list(dn) == dims
    for i, b in enumerate(blocks):
        o = int(dsh[i])
        ref = ddata[o:o + dims[i] ** 2].reshape(dims[i], dims[i])
        assert np.abs(b - ref).max(initial=0) < 1e-13, (name, i)
    assert abs(res["error"] - err_ref) < 1e-12, (name, res["error"], err_ref)
    sp = np.array(res["spectra"])
    assert sp.shape == z[tag + "_spectra"].shape
    assert np.abs(sp ** 2 - z[tag + "_spectra"] ** 2).max(initial=0) < 1e-13
    # rotation structure; every root's factor has the structure of root 0
    rq, rnb, rnk, _, rdata = ref_part(z, tag, "left")
    assert [int(x) for x in rq] == list(res["rot"][0])
    assert list(rnb) == list(res["rot"][1]) and list(rnk) == list(res["rot"][2])
    wq, wnb, wnk, wsh, wdata = ref_part(z, tag, "right")
    per_root = int(wsh[-1]) + int(wnb[-1]) * int(wnk[-1]) if len(wsh) else 0
    assert wdata.size == 2 * per_root
    lref = blocks_of(rq, rnb, rnk, _, rdata)
    lo = {int(a): d for a, d in zip(res["rot"][0], res["rot"][3])}
    for r in range(2):
        mq, mnb, mnk, mdata = res["wav"][r]
        assert [int(x) for x in wq] == mq, (name, r)
        assert list(wnb) == mnb and list(wnk) == mnk
        rref = wav_blocks((wq, wnb, wnk, wsh),
                          wdata[r * per_root:(r + 1) * per_root])
        ro = {int(a): d for a, d in zip(mq, mdata)}
        a = recon(sym, info, lref, rref, True)
        b = recon(sym, info, lo, ro, True)
        assert a.keys() == b.keys()
        for q in a:
            assert np.abs(a[q] - b[q]).max(initial=0) < 1e-8, (name, r, q)
        # each root's factor is normalised on its own
        assert abs(sum(float((x * x).sum()) for x in ro.values()) - 1) < 1e-12


def test_oracle_multi_split_reduces_to_single_root():
    """weights (1, 0) give the single-root split of the first root"""
    z, sym, info = load("hub8_sz")
    psi = z["psi_dav"]
    k = 40
    one = T.split_density_matrix(sym, info, psi, k, 1e-14, True, True, 0)
    two = T.multi_split_density_matrix(sym, info, [psi, z["sM_roots"]],
                                       [1.0, 0.0], k, 1e-14, True, True, 0)
    assert abs(one["error"] - two["error"]) < 1e-15
    for a, b in zip(one["dm"][2], two["dm"][2]):
        assert np.abs(a - b).max(initial=0) < 1e-15
    lo, ro = oracle_parts(one)
    wq, _, _, wd = two["wav"][0]
    for q, d in zip(wq, wd):
        assert np.abs(ro[q] - d).max() < 1e-12
