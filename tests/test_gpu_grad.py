"""GPU partial expectations (the backward of the two-site energy) against the
oracle and through torch.autograd.

* dE/dL, dE/dR of E = <v|H|c> against the oracle's orc_grad_ops (itself
  pinned by linearity and Euler's identity, tests/test_oracle.py) on every
  reference fixture and on a synthetic SU(2) site split into scratch chunks:
  1e-12 relative (2-norm);
* several right-hand sides sum;
* Energy.apply(psi, L, R): autograd gradients against central finite
  differences.  E is quadratic in psi and linear in L and R, so a central
  difference is exact up to rounding: 1e-8 relative per direction.  The
  n=26, M=1000 middle site uses 32 seeded directions with step 1e-5.
"""
import os

import numpy as np
import pytest
import torch

from conftest import FIXTURES, fixture_path

from b2g import synth
from b2g.effham import EffectiveHamiltonian, Energy
from b2g.problem import Problem
import b2g_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def dev(x):
    return torch.tensor(np.ascontiguousarray(x), dtype=torch.float64,
                        device="cuda:0")


@pytest.mark.parametrize("name", FIXTURES)
def test_grad_ops_match_oracle(gpu_ctx, name):
    print("case", name, flush=True)
    pb = Problem.from_npz(fixture_path(name))
    h = EffectiveHamiltonian(pb, gpu_ctx)
    rng = np.random.default_rng(3)
    v = rng.standard_normal(h.bra_len)
    c = rng.standard_normal(h.ket_len)
    gl, gr = h.grad_ops(dev(v), dev(c))
    ol, orr = O.Oracle(pb).grad_ops(v, c)
    assert rel(gl.cpu().numpy(), ol) < 1e-12, name
    assert rel(gr.cpu().numpy(), orr) < 1e-12, name
    h.deallocate()


def test_grad_ops_multi_rhs_and_chunks(gpu_ctx):
    pb = synth.middle_site(n_orb=16, m=500, max_dim=200, seed=3)
    rng = np.random.default_rng(8)
    old = os.environ.get("B2G_W_CAP")
    os.environ["B2G_W_CAP"] = "150000"  # several gradient scratch chunks
    try:
        h = EffectiveHamiltonian(pb, gpu_ctx)
    finally:
        if old is None:
            os.environ.pop("B2G_W_CAP")
        else:
            os.environ["B2G_W_CAP"] = old
    orc = O.Oracle(pb)
    v = rng.standard_normal((2, h.bra_len))
    c = rng.standard_normal((2, h.ket_len))
    gl, gr = h.grad_ops(dev(v), dev(c))
    ol = np.zeros_like(gl.cpu().numpy())
    orr = np.zeros_like(gr.cpu().numpy())
    for r in range(2):
        a, b = orc.grad_ops(v[r], c[r])
        ol += a
        orr += b
    assert rel(gl.cpu().numpy(), ol) < 1e-12
    assert rel(gr.cpu().numpy(), orr) < 1e-12
    h.deallocate()


def _fd_check(h, psi, lt, rt, ndir, step, seed):
    psi = psi.clone().requires_grad_(True)
    lv = lt.clone().requires_grad_(True)
    rv = rt.clone().requires_grad_(True)
    e = Energy.apply(psi, lv, rv, h)
    e.backward()
    g = torch.Generator(device="cuda:0").manual_seed(seed)
    worst = 0.0
    for k in range(ndir):
        which = k % 3
        base = (psi, lv, rv)[which].detach()
        d = torch.randn(base.shape, generator=g, dtype=torch.float64,
                        device="cuda:0")
        args_p = [psi.detach(), lv.detach(), rv.detach()]
        args_m = [psi.detach(), lv.detach(), rv.detach()]
        args_p[which] = base + step * d
        args_m[which] = base - step * d
        with torch.no_grad():
            ep = float(Energy.apply(*args_p, h))
            em = float(Energy.apply(*args_m, h))
        fd = (ep - em) / (2 * step)
        an = float(torch.dot((psi, lv, rv)[which].grad, d))
        err = abs(fd - an) / max(abs(an), 1e-30)
        worst = max(worst, err)
        assert err < 1e-8, (k, which, fd, an)
    # restore the unperturbed numbers in the effective Hamiltonian
    h.set_operator_data(lt, rt)
    return worst


def test_autograd_energy_reference_fixture(gpu_ctx):
    pb = Problem.from_npz(fixture_path("n2_su2"))
    h = EffectiveHamiltonian(pb, gpu_ctx)
    psi = dev(pb.golden["psi"])
    lt, rt = h.operator_data()
    # dE/dpsi = 2 H psi against the reference's own H*psi
    p = psi.clone().requires_grad_(True)
    Energy.apply(p, lt, rt, h).backward()
    assert rel(p.grad.cpu().numpy(), 2 * pb.golden["hpsi"]) < 1e-12
    _fd_check(h, psi, lt, rt, ndir=9, step=1e-5, seed=1)
    h.deallocate()


def test_autograd_energy_n26_m1000_middle_site(gpu_ctx):
    # backward at the n=26, M=1000 SU(2) middle site: 32 seeded directions
    pb = synth.middle_site(n_orb=26, m=1000, max_dim=300, seed=26)
    print("n26 |psi| %d" % pb.ket_len, flush=True)
    h = EffectiveHamiltonian(pb, gpu_ctx)
    g = torch.Generator(device="cuda:0").manual_seed(5)
    psi = torch.randn(h.ket_len, generator=g, dtype=torch.float64,
                      device="cuda:0")
    psi /= torch.linalg.norm(psi)
    lt, rt = h.operator_data()
    worst = _fd_check(h, psi, lt, rt, ndir=32, step=1e-5, seed=2)
    print("n26 worst relative FD error %.3g" % worst)
    h.deallocate()
