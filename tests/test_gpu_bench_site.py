"""GPU parity on the benchmarked workload itself: the middle-site effective
Hamiltonian of the n=40-orbital, M=3000 SU(2) problem (seed 1234, 162 OpSum
strings, 47,204 recorded pairs, blocks up to 600, |psi| = 2,341,066, split-K
active), 4 right-hand sides, exactly as bench.py builds it (operator data
generated on the device).

* FP64 H*psi against the oracle on a seeded set of output blocks that
  includes the block with the most multiply-adds, a 600-row block and the
  deepest (split-K) block, for the first and the last right-hand side:
  1e-12 relative; FP32 2e-6; split-BF16 1e-5.
* Size-independent properties of the full product: <x|H y> = <y|H x>
  (H is symmetric) to 1e-12 relative, and the 4-RHS product bitwise equal to
  four 1-RHS products (outputs are tile-owned, reductions in fixed order).
* The split-BF16 product of effective Hamiltonians rebuilt from the same
  problem (packed plan derived again on the host workers) bitwise equal.
* The diagonal against the oracle's, 1e-12.
* Four state-averaged roots, graph-captured Davidson: mixed precision
  (FP32 and split-BF16 products + FP64 refinement) against FP64, |dE| <=
  1e-6 Eh for every root.
"""
import os

import numpy as np
import pytest
import torch

from b2g import synth
from b2g.effham import EffectiveHamiltonian
import b2g_oracle as O

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def site(gpu_ctx):
    print("bench site: n=40 M=3000 max_dim 600 seed 1234 pair_scale 4",
          flush=True)
    pb = synth.middle_site(n_orb=40, m=3000, max_dim=600, seed=1234,
                           pair_scale=4, data="device")
    ldev, rdev = synth.fill_device(pb, "cuda:0")
    torch.cuda.synchronize()
    h = EffectiveHamiltonian(pb, gpu_ctx, fp32=True,
                             device_data=(ldev, rdev))
    assert h.n_pairs == 47204 and h.ket_len == 2341066
    h.device_data = (ldev, rdev)
    pb.lops.data = ldev.cpu().numpy()
    pb.rops.data = rdev.cpu().numpy()
    orc = O.Oracle(pb)
    g = torch.Generator(device="cuda:0").manual_seed(1234)
    x = torch.randn((4, h.ket_len), generator=g, dtype=torch.float64,
                    device="cuda:0")
    # output blocks: most multiply-adds, a 600-row block, the deepest
    # (largest summed contraction depth, split into partial tiles), and
    # three seeded others
    macs = orc.block_macs().astype(np.float64)
    nb = pb.bra.nbra.astype(np.float64)
    nk = pb.bra.nket.astype(np.float64)
    depth = macs / np.maximum(nb * nk, 1.0)
    sel = [int(np.argmax(macs)), int(np.argmax(nb)), int(np.argmax(depth))]
    rng = np.random.default_rng(5)
    sel += [int(b) for b in rng.choice(len(nb), 3, replace=False)]
    sel = sorted(set(sel))
    print("blocks", sel, "dims", [(int(nb[b]), int(nk[b])) for b in sel],
          "depth", [int(depth[b]) for b in sel], flush=True)
    mask = np.zeros(h.bra_len, bool)
    for b in sel:
        s0 = int(pb.bra.shift[b])
        mask[s0:s0 + int(nb[b] * nk[b])] = True
    ref = {}
    xh = x.cpu().numpy()
    for r in (0, 3):
        ref[r], _ = orc.hpsi_blocks(xh[r], sel, THREADS)
    yield pb, h, orc, x, sel, mask, ref
    h.deallocate()


@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("f32", 2e-6),
                                           ("bf16x3", 1e-5)])
def test_selected_blocks_against_oracle(site, precision, tol):
    pb, h, orc, x, sel, mask, ref = site
    y = h.multiply(x, precision=precision).cpu().numpy()
    for r in (0, 3):
        err = rel(y[r][mask], ref[r][mask])
        print(precision, r, "rel err %.3g" % err, flush=True)
        assert err < tol, (precision, r, err)


def test_symmetry(site):
    pb, h, orc, x, sel, mask, ref = site
    y = h.multiply(x[:2]).cpu().numpy()
    xh = x[:2].cpu().numpy()
    a, b = float(xh[0] @ y[1]), float(xh[1] @ y[0])
    scale = np.linalg.norm(xh[0]) * np.linalg.norm(y[1])
    assert abs(a - b) < 1e-12 * scale, (a, b)


def test_multi_rhs_equals_single_rhs_bitwise(site):
    pb, h, orc, x, sel, mask, ref = site
    y4 = h.multiply(x).cpu().numpy()
    for r in range(4):
        y1 = h.multiply(x[r]).cpu().numpy()
        assert np.array_equal(y1, y4[r]), r


def test_packed_plan_rebuild_bitwise(site, gpu_ctx):
    """The split-BF16 packed plan is derived by host worker threads (lists,
    tiles and task order as separate jobs): a second effective Hamiltonian
    built from the same problem, and a third after the workers have built
    other plans, give bitwise the same split-BF16 product."""
    pb, h, orc, x, sel, mask, ref = site
    y0 = h.multiply(x[:2], precision="bf16x3").cpu().numpy()
    ldata, rdata = h.device_data
    for _ in range(2):
        h2 = EffectiveHamiltonian(pb, gpu_ctx, fp32=False,
                                  device_data=(ldata, rdata))
        y2 = h2.multiply(x[:2], precision="bf16x3").cpu().numpy()
        h2.deallocate()
        assert np.array_equal(y0, y2)


def test_diagonal_against_oracle(site):
    pb, h, orc, x, sel, mask, ref = site
    assert rel(h.diag.cpu().numpy(), orc.diag()) < 1e-12


def test_four_roots_mixed_against_fp64(site):
    pb, h, orc, x, sel, mask, ref = site
    d = h.diag.cpu().numpy()
    idx = np.argsort(d)[:4]
    rng = np.random.default_rng(8)
    x0 = 1e-4 * rng.standard_normal((4, h.ket_len))
    for i in range(4):
        x0[i, idx[i]] += 1.0
    out = {}
    for prec in ("f64", "f32", "bf16x3"):
        xv = torch.tensor(x0, dtype=torch.float64, device="cuda:0")
        e, nd, nflop, td = h.eigs(xv, conv_thrd=1e-8, max_iter=200,
                                  precision=prec, refine=True)
        out[prec] = (np.array(e), xv.cpu().numpy())
        print(prec, e, nd, "%.3f s" % td, flush=True)
    e64, v64 = out["f64"]
    for prec in ("f32", "bf16x3"):
        e, v = out[prec]
        assert np.abs(e - e64).max() <= 1e-6, (prec, e, e64)
        ov = np.abs(np.sum(v * v64, axis=1))
        print(prec, "infidelity", 1 - ov, flush=True)
