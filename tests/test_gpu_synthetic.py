"""GPU H*psi on seeded synthetic middle-site problems against the oracle:
every tile configuration of the grouped GEMM (32x32, 64x64, 128x128 forced
through B2G_TILE, and the automatic choice), the scratch split into several
chunks, split-K partial tiles (B2G_KSPLIT), 4 right-hand sides, FP32 and
split-BF16 three-term (every tile shape).  Blocks range from 1x1 up to 200x200, so
edge tiles, K remainders and single-element blocks all occur.
Tolerances: FP64 1e-12 relative (2-norm), FP32 2e-6, split-BF16 1e-4."""
import os

import numpy as np
import pytest
import torch

from b2g import synth
from b2g.effham import EffectiveHamiltonian
import b2g_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def problem():
    pb = synth.middle_site(n_orb=16, m=500, max_dim=200, seed=3)
    orc = O.Oracle(pb)
    rng = np.random.default_rng(9)
    xs = rng.standard_normal((4, pb.ket_len))
    ys = np.stack([orc.hpsi(x, nthreads=8) for x in xs])
    return pb, xs, ys


def run(gpu_ctx, pb, xs, env, precision="f64"):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        h = EffectiveHamiltonian(pb, gpu_ctx)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    x = torch.tensor(xs, dtype=torch.float64, device="cuda:0")
    y = h.multiply(x, precision=precision).cpu().numpy()
    h.deallocate()
    return y


TILES = ["32", "64", "128", "32x64", "64x32", "32x128", "128x32", "64x128",
         "128x64", "auto"]


@pytest.mark.parametrize("tile", TILES)
def test_tile_configurations(gpu_ctx, problem, tile):
    pb, xs, ys = problem
    print("tile", tile, flush=True)
    env = {} if tile == "auto" else {"B2G_TILE": tile}
    y = run(gpu_ctx, pb, xs, env)
    for r in range(4):
        assert rel(y[r], ys[r]) < 1e-12, (tile, r)


def test_scratch_chunks(gpu_ctx, problem):
    pb, xs, ys = problem
    # about a tenth of the W scratch per chunk -> many chunks that accumulate
    y = run(gpu_ctx, pb, xs, {"B2G_W_CAP": "200000"})
    for r in range(4):
        assert rel(y[r], ys[r]) < 1e-12, r


@pytest.mark.parametrize("ksplit", ["1", "64", "1000"])
def test_split_k(gpu_ctx, problem, ksplit):
    # deep output blocks cut into partial tiles plus a fixed-order reduction;
    # with several scratch chunks the reductions accumulate into y
    pb, xs, ys = problem
    print("ksplit", ksplit, flush=True)
    for env in ({"B2G_KSPLIT": ksplit},
                {"B2G_KSPLIT": ksplit, "B2G_W_CAP": "200000"}):
        y = run(gpu_ctx, pb, xs, env)
        for r in range(4):
            assert rel(y[r], ys[r]) < 1e-12, (env, r)
    y = run(gpu_ctx, pb, xs, {"B2G_KSPLIT": ksplit}, precision="f32")
    for r in range(4):
        assert rel(y[r], ys[r]) < 2e-6, r


@pytest.mark.parametrize("tile", TILES)
def test_bf16x3_tile_configurations(gpu_ctx, problem, tile):
    # split-BF16 three-term kernels, every tile shape; 1e-4 relative
    pb, xs, ys = problem
    print("bf16x3 tile", tile, flush=True)
    env = {} if tile == "auto" else {"B2G_TILE": tile}
    y = run(gpu_ctx, pb, xs, env, precision="bf16x3")
    for r in range(4):
        assert rel(y[r], ys[r]) < 1e-4, (tile, r, rel(y[r], ys[r]))


def test_fp32(gpu_ctx, problem):
    pb, xs, ys = problem
    y = run(gpu_ctx, pb, xs, {}, precision="f32")
    for r in range(4):
        assert rel(y[r], ys[r]) < 2e-6, r
