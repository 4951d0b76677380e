"""GPU parity of the device truncation (b2g_split_density_matrix) and its
batched eigensolver (b2g_syev_batched).

Eigensolver: random symmetric matrices of orders around the 32-column block
edges (1 .. 333), indefinite, positive semi-definite of low rank (the shape
of a density matrix) and with a repeated eigenvalue; eigenvalues against
numpy.linalg.eigh to 1e-12 * ||A||_2, residual ||A v - w v|| and
orthonormality to 1e-11 * ||A||_2; eigenvectors of simple eigenvalues equal
to the oracle's signed ones to 1e-9.

Split: on every golden fixture and truncation case (three without noise,
two with block2's wavefunction / density-matrix noise drawn from its
generator seeded with 4321), and at every site dumped along block2's sweeps
(wavefunction noise, density-matrix noise, and density-matrix noise with
every state kept in both directions), against the reference's own dumps (the oracle is pinned to the same dumps on the CPU):
density matrix to 1e-13, kept states per sector exactly (up to states
whose eigenvalue lies within 1e-13 of the cutoff, where FP64 eigensolvers
may fall on either side; then the structure and reconstruction
comparisons below are skipped), error to 1e-12, squared
spectra to 1e-13 (all absolute; the density matrix has trace 1); the
kept vectors span an invariant subspace of the reference's density matrix
(||D L - L (L^T D L)|| <= 1e-12) and are orthonormal (1e-12); the
wavefunction factor equals L^T psi / ||L^T psi|| (or psi R^T / ||.||)
recomputed on the host (1e-12); the truncated wavefunction left * right
equals the reference's to 1e-7: states whose density-matrix eigenvalues lie
at the rounding level of ||D|| (1e-14 .. 1e-12, kept above the 1e-14
cutoff) are determined only up to rotations among themselves by any
eigensolver, and carry weight sqrt(lambda) ~ 1e-7 in the wavefunction; and
on a synthetic n=26-orbital
middle-site wavefunction against the oracle: density matrix, eigenvalues
and error to 1e-13 / 1e-12, structure exactly, and the same invariant-
subspace and factor checks.
"""
import json

import numpy as np
import pytest
import torch

import truncate_oracle as T
from conftest import FIXTURES, DELAYED_FIXTURES, fixture_path

from b2g import truncate as TR
from sweep_common import (SWEEP_SPLIT_CASES, load_sweep_split,
                          sweep_split_fixtures)

pytestmark = pytest.mark.gpu

CASES = ["sR", "sL", "sK", "sW", "sD"]


def settings(z, tag):
    """err, k, cutoff, trunc_type, trace_right, noise, noise_type, seed"""
    st = [float(x) for x in z[tag + "_settings"]]
    return st + [0.0, 0.0, 0.0][:8 - len(st)]


def dev(x):
    return torch.tensor(np.ascontiguousarray(x), dtype=torch.float64,
                        device="cuda:0")


def sym_mats():
    rng = np.random.default_rng(7)
    mats = []
    for n in [1, 2, 5, 31, 32, 33, 63, 64, 65, 100, 129, 200, 333]:
        a = rng.standard_normal((n, n))
        mats.append(("indef%d" % n, (a + a.T) / 2))
    for n, r in [(40, 7), (150, 30), (257, 64)]:
        x = rng.standard_normal((n, r))
        mats.append(("psd%d_r%d" % (n, r), x @ x.T / n))
    q, _ = np.linalg.qr(rng.standard_normal((90, 90)))
    w = np.concatenate([np.full(10, 0.5), rng.uniform(-1, 1, 80)])
    mats.append(("degen90", (q * w) @ q.T))
    return mats


def test_syev_batched_matches_numpy(gpu_ctx):
    mats = sym_mats()
    ws, vs, sweeps = TR.eigh_batched(gpu_ctx, [dev(m) for _, m in mats])
    print("block Jacobi sweeps", sweeps)
    for (name, a), w, v in zip(mats, ws, vs):
        w = w.cpu().numpy()
        v = v.cpu().numpy()
        wn = np.linalg.eigvalsh(a)
        nrm = max(np.abs(wn).max(), 1e-300)
        assert np.abs(w - wn).max() < 1e-12 * nrm, name
        assert np.abs(a @ v.T - v.T * w).max() < 1e-11 * nrm, name
        assert np.abs(v @ v.T - np.eye(len(w))).max() < 1e-11, name
        # signed eigenvectors of simple eigenvalues match the oracle's
        wo, vo = T.signed_eigh(a)
        gap = np.diff(wo)
        for j in range(len(w)):
            lo = gap[j - 1] if j > 0 else np.inf
            hi = gap[j] if j < len(gap) else np.inf
            if min(lo, hi) > 1e-6 * nrm:
                assert np.abs(v[j] - vo[j]).max() < 1e-9, (name, j)


def test_syev_batched_repeatable(gpu_ctx):
    mats = [dev(m) for _, m in sym_mats()]
    w1, v1, _ = TR.eigh_batched(gpu_ctx, mats)
    w2, v2, _ = TR.eigh_batched(gpu_ctx, mats)
    for a, b in zip(w1 + v1, w2 + v2):
        assert torch.equal(a, b)


def load(name):
    z = np.load(fixture_path(name))
    meta = json.loads(bytes(z["meta_json"]).decode())
    sym = T.SU2 if meta["sym"] == "su2" else T.SZ
    info = TR.BlockInfo(z["ket_quanta"], z["ket_nbra"], z["ket_nket"],
                        z["ket_shift"], int(z["ket_dq"][0]))
    return z, sym, info


def part_blocks(bi, data):
    d = data.cpu().numpy()
    return {int(q): bi.block(d, i) for i, q in enumerate(bi.quanta)}


def recon(sym, info, left, right, trace_right):
    dq = info.dq
    lab = {int(q): (T.q_get_bra(sym, int(q), dq) if trace_right
                    else T.q_neg(sym, T.q_get_ket(sym, int(q))))
           for q in info.quanta}
    rot, wav = (left, right) if trace_right else (right, left)
    return {q: (rot[lab[q]] @ w if trace_right else w @ rot[lab[q]])
            for q, w in wav.items()}


def check_factors(sym, info, psi, mine, dref, tr, name):
    """kept vectors: an orthonormal invariant subspace of the density
    matrix dref (dict quantum -> block); wavefunction factor = L^T psi
    (psi R^T), normalised"""
    rot = mine["left"] if tr else {q: v.T for q, v in mine["right"].items()}
    for q, lv in rot.items():
        d = dref[q]
        assert np.abs(lv.T @ lv - np.eye(lv.shape[1])).max() < 1e-12, (name, q)
        assert np.abs(d @ lv - lv @ (lv.T @ d @ lv)).max() < 1e-12, (name, q)
    dq = info.dq
    wav = mine["right"] if tr else mine["left"]
    raw = {}
    for i, q in enumerate(info.quanta):
        q = int(q)
        if q not in wav:
            continue
        lab = (T.q_get_bra(sym, q, dq) if tr
               else T.q_neg(sym, T.q_get_ket(sym, q)))
        x = info.block(psi, i)
        raw[q] = rot[lab].T @ x if tr else x @ rot[lab]
    nrm = np.sqrt(sum(float((v * v).sum()) for v in raw.values()))
    assert raw.keys() == wav.keys()
    for q in raw:
        assert np.abs(raw[q] / nrm - wav[q]).max(initial=0) < 1e-12, (name, q)


def kept_matches(z, tag, s, sym, tt, tr, cutoff, name):
    """Kept states per sector equal to the reference's, except for states
    whose (reduced) eigenvalue lies within 1e-13 of the cutoff, where FP64
    eigensolvers (LAPACK's absolute error is n eps ||D|| ~ 1e-14 here) may
    fall on either side.  Returns whether they are equal everywhere."""
    sp = z[tag + "_spectra"]
    dmi = s.dm[0]
    rot_part = "left" if tr else "right"
    rp = tag + "_" + rot_part
    kept_ref = {int(q): int(m) for q, m in zip(
        z[rp + "_quanta"], z[rp + ("_nket" if tr else "_nbra")])}
    rbi = s.left[0] if tr else s.right[0]
    kept = {int(q): int(m) for q, m in zip(
        rbi.quanta, rbi.nket if tr else rbi.nbra)}
    amb, o = {}, 0
    for i, q in enumerate(dmi.quanta):
        n = int(dmi.nbra[i])
        lam = sp[o:o + n] ** 2
        if tt & 1:
            lam = lam / T.multiplicity(sym, int(q))
        amb[int(q)] = int((np.abs(lam - cutoff) < 1e-13).sum())
        o += n
    strict = True
    for q in set(kept) | set(kept_ref):
        d = abs(kept.get(q, 0) - kept_ref.get(q, 0))
        assert d <= amb.get(q, 0), (name, tag, q, kept.get(q), kept_ref.get(q))
        strict = strict and d == 0
    return strict


@pytest.mark.parametrize("name", FIXTURES + DELAYED_FIXTURES)
@pytest.mark.parametrize("tag", CASES)
def test_split_matches_reference(gpu_ctx, name, tag):
    z, sym, info = load(name)
    if tag + "_settings" not in z.files:
        pytest.skip("fixture without truncation dump")
    check_device_split(gpu_ctx, z, sym, info, name, tag)


@pytest.mark.parametrize("case,env", sweep_split_fixtures())
@pytest.mark.parametrize("tag", SWEEP_SPLIT_CASES)
def test_split_matches_sweep_dumps(gpu_ctx, case, env, tag):
    z, sym, inf = load_sweep_split(case, env)
    info = TR.BlockInfo(inf["quanta"], inf["nbra"], inf["nket"],
                        inf["shift"], inf["dq"])
    check_device_split(gpu_ctx, z, sym, info, case + "/" + env, tag)


def check_device_split(gpu_ctx, z, sym, info, name, tag):
    err_ref, k, cutoff, tt, tr, noise, nt, seed = settings(z, tag)
    print("case", name, tag, flush=True)
    s = TR.split_density_matrix(gpu_ctx, sym, info, dev(z["psi_dav"]),
                                int(k), cutoff, bool(tr), True, int(tt),
                                noise=noise, noise_type=int(nt),
                                seed=int(seed))
    # density matrix (with density-matrix noise block2's is asymmetric and
    # its eigensolver reads the lower triangle; the device keeps that
    # triangle, mirrored)
    dmi, dmd = s.dm
    p = tag + "_dm"
    assert np.array_equal(dmi.quanta, z[p + "_quanta"].astype(np.uint64))
    assert np.array_equal(dmi.nbra, z[p + "_nbra"])
    dref = {int(q): T.lower_symmetric(dmi.block(z[p + "_data"], i))
            for i, q in enumerate(dmi.quanta)}
    dd = dmd.cpu().numpy()
    for i, q in enumerate(dmi.quanta):
        assert np.abs(dmi.block(dd, i) - dref[int(q)]).max(initial=0) < 1e-13
    # error, spectra
    assert abs(s.error - err_ref) < 1e-12, (name, tag, s.error, err_ref)
    sp = z[tag + "_spectra"]
    assert s.spectra.shape == sp.shape
    assert np.abs(s.spectra ** 2 - sp ** 2).max(initial=0) < 1e-13
    strict = kept_matches(z, tag, s, sym, int(tt), bool(tr), cutoff, name)
    mine = {"left": part_blocks(*s.left), "right": part_blocks(*s.right)}
    check_factors(sym, info, z["psi_dav"], mine, dref, bool(tr), name)
    if not strict:
        return
    # factor structure
    for part, (bi, _) in (("left", s.left), ("right", s.right)):
        p = tag + "_" + part
        assert np.array_equal(bi.quanta, z[p + "_quanta"].astype(np.uint64))
        assert np.array_equal(bi.nbra, z[p + "_nbra"])
        assert np.array_equal(bi.nket, z[p + "_nket"])
    # truncated wavefunction, sign-free
    ref = {}
    for part in ("left", "right"):
        p = tag + "_" + part
        rb = TR.BlockInfo(z[p + "_quanta"], z[p + "_nbra"], z[p + "_nket"],
                          z[p + "_shift"])
        ref[part] = {int(q): rb.block(z[p + "_data"], i)
                     for i, q in enumerate(rb.quanta)}
    a = recon(sym, info, ref["left"], ref["right"], bool(tr))
    b = recon(sym, info, mine["left"], mine["right"], bool(tr))
    assert a.keys() == b.keys()
    for q in a:
        assert np.abs(a[q] - b[q]).max(initial=0) < 1e-7, (name, tag, q)


@pytest.mark.parametrize("trace_right", [True, False])
def test_split_synthetic_matches_oracle(gpu_ctx, trace_right):
    """A synthetic n=26 SU(2) middle-site wavefunction (random, normalised;
    sector histogram of the M=1000 generator, dimensions to ~300)."""
    from b2g import synth
    pb = synth.middle_site(n_orb=26, m=1000, max_dim=300, seed=11,
                           data="device")
    o, n = int(pb.ket.off[0]), int(pb.ket.n[0])
    info = TR.BlockInfo(pb.ket.quanta[o:o + n], pb.ket.nbra[o:o + n],
                        pb.ket.nket[o:o + n], pb.ket.shift[o:o + n],
                        int(pb.ket.dq[0]))
    rng = np.random.default_rng(3)
    psi = rng.standard_normal(info.total)
    psi /= np.linalg.norm(psi)
    sym = T.SU2
    k = 1000
    s = TR.split_density_matrix(gpu_ctx, sym, info, dev(psi), k, 1e-14,
                                trace_right, True, 0)
    oi = dict(quanta=info.quanta, nbra=info.nbra, nket=info.nket,
              shift=info.shift, dq=info.dq)
    res = T.split_density_matrix(sym, oi, psi, k, 1e-14, trace_right, True, 0)
    assert abs(s.error - res["error"]) < 1e-12
    wo = np.concatenate([np.asarray(x) for x in res["eigvals"]])
    assert np.abs(s.eigvals - wo).max() < 1e-13
    uq, dims, blocks = res["dm"]
    dmi, dmd = s.dm
    assert [int(x) for x in dmi.quanta] == uq
    dd = dmd.cpu().numpy()
    for i, blk in enumerate(blocks):
        assert np.abs(dmi.block(dd, i) - blk).max(initial=0) < 1e-13
    for part, mine in (("left", s.left), ("right", s.right)):
        q, nb, nk, data = res[part]
        bi, d = mine
        assert [int(x) for x in bi.quanta] == list(q)
        assert list(bi.nbra) == list(nb) and list(bi.nket) == list(nk)
    mine = {"left": part_blocks(*s.left), "right": part_blocks(*s.right)}
    check_factors(sym, info, psi, mine, dict(zip(uq, blocks)), trace_right,
                  "synthetic")


def root_parts(s, r, tr):
    """(rotation blocks, root r's wavefunction-factor blocks) of a
    state-averaged split, as the left/right dict check_factors takes"""
    wav = s.right if tr else s.left
    rot = s.left if tr else s.right
    w = {"wav": part_blocks(wav[0], wav[1][r]), "rot": part_blocks(*rot)}
    return ({"left": w["rot"], "right": w["wav"]} if tr
            else {"left": w["wav"], "right": w["rot"]})


@pytest.mark.parametrize("name", FIXTURES + DELAYED_FIXTURES)
@pytest.mark.parametrize("tag", ["sM", "sN", "sQ"])
def test_multi_split_matches_reference(gpu_ctx, name, tag):
    """State-averaged split of two roots (the Davidson vector and a
    normalised random vector, weights 0.7 / 0.3) against the reference's
    density_matrix_with_multi_target + multi_split_density_matrix: the
    same checks as the single-root split, with the factor checks for each
    root; without noise (sM) and with block2's wavefunction (sN) and
    density-matrix (sQ) noise 1e-6, one draw per root from its generator
    seeded with 4321."""
    z, sym, info = load(name)
    if tag + "_settings" not in z.files:
        pytest.skip("fixture without multi-root dump")
    err_ref, k, cutoff, tt, tr, noise, nt, seed = settings(z, tag)
    print("case", name, tag, flush=True)
    roots = [z["psi_dav"], z[tag + "_roots"]]
    s = TR.split_density_matrix(gpu_ctx, sym, info, dev(np.stack(roots)),
                                int(k), cutoff, bool(tr), True, int(tt),
                                weights=[0.7, 0.3], noise=noise,
                                noise_type=int(nt), seed=int(seed))
    dmi, dmd = s.dm
    p = tag + "_dm"
    assert np.array_equal(dmi.quanta, z[p + "_quanta"].astype(np.uint64))
    assert np.array_equal(dmi.nbra, z[p + "_nbra"])
    dref = {int(q): T.lower_symmetric(dmi.block(z[p + "_data"], i))
            for i, q in enumerate(dmi.quanta)}
    dd = dmd.cpu().numpy()
    for i, q in enumerate(dmi.quanta):
        assert np.abs(dmi.block(dd, i) - dref[int(q)]).max(initial=0) < 1e-13
    assert abs(s.error - err_ref) < 1e-12, (name, s.error, err_ref)
    sp = z[tag + "_spectra"]
    assert s.spectra.shape == sp.shape
    assert np.abs(s.spectra ** 2 - sp ** 2).max(initial=0) < 1e-13
    strict = kept_matches(z, tag, s, sym, int(tt), bool(tr), cutoff, name)
    wav = s.right if tr else s.left
    assert wav[1].shape == (2, wav[0].total)
    for r in range(2):
        check_factors(sym, info, roots[r], root_parts(s, r, bool(tr)), dref,
                      bool(tr), (name, r))
    if not strict:
        return
    for part, (bi, _) in (("left", s.left), ("right", s.right)):
        p = tag + "_" + part
        assert np.array_equal(bi.quanta, z[p + "_quanta"].astype(np.uint64))
        assert np.array_equal(bi.nbra, z[p + "_nbra"])
        assert np.array_equal(bi.nket, z[p + "_nket"])
    lb = TR.BlockInfo(z[tag + "_left_quanta"], z[tag + "_left_nbra"],
                      z[tag + "_left_nket"], z[tag + "_left_shift"])
    rb = TR.BlockInfo(z[tag + "_right_quanta"], z[tag + "_right_nbra"],
                      z[tag + "_right_nket"], z[tag + "_right_shift"])
    lref = {int(q): lb.block(z[tag + "_left_data"], i)
            for i, q in enumerate(lb.quanta)}
    rdata = z[tag + "_right_data"].reshape(2, -1)
    for r in range(2):
        rref = {int(q): rb.block(rdata[r], i) for i, q in enumerate(rb.quanta)}
        mine = root_parts(s, r, True)
        a = recon(sym, info, lref, rref, True)
        b = recon(sym, info, mine["left"], mine["right"], True)
        assert a.keys() == b.keys()
        for q in a:
            assert np.abs(a[q] - b[q]).max(initial=0) < 1e-7, (name, r, q)


@pytest.mark.parametrize("trace_right", [True, False])
def test_multi_split_synthetic_matches_oracle(gpu_ctx, trace_right):
    """Three random roots of the synthetic n=26 middle-site structure,
    weights 0.5 / 0.3 / 0.2, both directions, against the oracle."""
    from b2g import synth
    pb = synth.middle_site(n_orb=26, m=1000, max_dim=300, seed=11,
                           data="device")
    o, n = int(pb.ket.off[0]), int(pb.ket.n[0])
    info = TR.BlockInfo(pb.ket.quanta[o:o + n], pb.ket.nbra[o:o + n],
                        pb.ket.nket[o:o + n], pb.ket.shift[o:o + n],
                        int(pb.ket.dq[0]))
    rng = np.random.default_rng(5)
    psis = rng.standard_normal((3, info.total))
    psis /= np.linalg.norm(psis, axis=1)[:, None]
    wts = [0.5, 0.3, 0.2]
    sym = T.SU2
    k = 800
    s = TR.split_density_matrix(gpu_ctx, sym, info, dev(psis), k, 1e-14,
                                trace_right, True, 0, weights=wts)
    oi = dict(quanta=info.quanta, nbra=info.nbra, nket=info.nket,
              shift=info.shift, dq=info.dq)
    res = T.multi_split_density_matrix(sym, oi, list(psis), wts, k, 1e-14,
                                       trace_right, True, 0)
    assert abs(s.error - res["error"]) < 1e-12
    wo = np.concatenate([np.asarray(x) for x in res["eigvals"]])
    assert np.abs(s.eigvals - wo).max() < 1e-13
    uq, dims, blocks = res["dm"]
    dmi, dmd = s.dm
    assert [int(x) for x in dmi.quanta] == uq
    dd = dmd.cpu().numpy()
    for i, blk in enumerate(blocks):
        assert np.abs(dmi.block(dd, i) - blk).max(initial=0) < 1e-13
    rot = s.left if trace_right else s.right
    assert [int(x) for x in rot[0].quanta] == list(res["rot"][0])
    assert list(rot[0].nbra) == list(res["rot"][1])
    assert list(rot[0].nket) == list(res["rot"][2])
    wav = s.right if trace_right else s.left
    for r in range(3):
        wq, wnb, wnk, _ = res["wav"][r]
        assert [int(x) for x in wav[0].quanta] == wq
        assert list(wav[0].nbra) == wnb and list(wav[0].nket) == wnk
        check_factors(sym, info, psis[r], root_parts(s, r, trace_right),
                      dict(zip(uq, blocks)), trace_right, ("synthetic", r))
