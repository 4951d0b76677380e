"""The allocation-free H*psi (b2g_hpsi_noalloc, scratch in the library's
fixed device array) against b2g_hpsi (plan-owned scratch) through the C
ABI: bit-identical results on two calls in a row, the second on a smaller
problem, in FP64 (one and four right-hand sides, overwrite and accumulate)
and FP32; a product whose scratch exceeds the array is refused before
anything runs, and so is split-BF16."""
import ctypes as C

import numpy as np
import pytest
import torch

from b2g import _lib as L
from b2g import synth
from b2g.effham import EffectiveHamiltonian

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def _products(ctx, pb, nrhs, precision, accumulate, seed):
    lib = L.lib()
    h = EffectiveHamiltonian(pb, ctx, fp32=(precision == "f32"))
    plan = h.plan
    rng = np.random.default_rng(seed)
    x = torch.tensor(rng.standard_normal((nrhs, pb.ket_len)),
                     dtype=torch.float64, device="cuda:0")
    y0 = torch.tensor(rng.standard_normal((nrhs, pb.bra_len)),
                      dtype=torch.float64, device="cuda:0")
    y_ref, y_new = y0.clone(), y0.clone()
    torch.cuda.synchronize()
    prec = L.PREC[precision]
    L.check(lib.b2g_hpsi(ctx.h, plan, x.data_ptr(), y_ref.data_ptr(), nrhs,
                         0.75, accumulate, prec))
    ctx.synchronize()
    L.check(lib.b2g_hpsi_noalloc(ctx.h, plan, x.data_ptr(), y_new.data_ptr(),
                                 nrhs, 0.75, accumulate, prec))
    ctx.synchronize()
    w, t1, t2 = C.c_uint64(), C.c_uint64(), C.c_uint64()
    L.check(lib.b2g_plan_stats(plan, C.byref(t1), C.byref(t2), C.byref(w)))
    h.deallocate()
    return y_ref.cpu().numpy(), y_new.cpu().numpy(), int(w.value)


@pytest.mark.parametrize("nrhs,precision,accumulate",
                         [(1, "f64", 0), (4, "f64", 1), (2, "f32", 0)])
def test_noalloc_matches_hpsi_on_two_calls(gpu_ctx, nrhs, precision,
                                           accumulate):
    # first a larger problem, then a smaller one: the second call must not
    # see anything the first left in the shared scratch
    for n_orb, m, dim, seed in [(12, 200, 80, 5), (8, 60, 30, 6)]:
        pb = synth.middle_site(n_orb=n_orb, m=m, max_dim=dim, seed=seed)
        print("noalloc", n_orb, m, nrhs, precision, accumulate, flush=True)
        ref, new, w = _products(gpu_ctx, pb, nrhs, precision, accumulate,
                                seed)
        assert w * nrhs * 8 < SCRATCH_BYTES, w
        assert np.isfinite(ref).all()
        assert np.array_equal(ref, new), np.abs(ref - new).max()


def test_noalloc_refuses_oversized_and_bf16(gpu_ctx):
    lib = L.lib()
    pb = synth.middle_site(n_orb=12, m=200, max_dim=80, seed=5)
    h = EffectiveHamiltonian(pb, gpu_ctx)
    w, t1, t2 = C.c_uint64(), C.c_uint64(), C.c_uint64()
    L.check(lib.b2g_plan_stats(h.plan, C.byref(t1), C.byref(t2), C.byref(w)))
    # enough right-hand sides that the W windows alone exceed the array;
    # refused on the host before any launch (the vectors are never read)
    nrhs = SCRATCH_BYTES // (8 * max(int(w.value), 1)) + 1
    assert nrhs <= 65535
    x = torch.zeros(pb.ket_len, dtype=torch.float64, device="cuda:0")
    y = torch.zeros(pb.bra_len, dtype=torch.float64, device="cuda:0")
    r = lib.b2g_hpsi_noalloc(gpu_ctx.h, h.plan, x.data_ptr(), y.data_ptr(),
                             nrhs, 1.0, 0, L.PREC["f64"])
    assert r != 0
    assert b"scratch" in lib.b2g_last_error()
    r = lib.b2g_hpsi_noalloc(gpu_ctx.h, h.plan, x.data_ptr(), y.data_ptr(),
                             1, 1.0, 0, L.PREC["bf16x3"])
    assert r != 0
    h.deallocate()
