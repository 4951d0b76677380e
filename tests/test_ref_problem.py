"""Block2's own H*psi and diagonal on problems handed to it as arrays
(oracle/ref_problem.py -> oracle/_ref/b2ref --load-problem, CPU, in this
container): the reference builds its SparseMatrixInfo / SparseMatrix /
OpSum objects and its EffectiveHamiltonian from them and runs its Tasked
replay.

* On seeded synthetic SU(2) middle sites (the generator of the benchmark
  workload) block2's H x equals the oracle's to 1e-12 relative, its
  diagonal the oracle's, and its multiply-add count the engine's plan
  count: the synthetic workload is a problem block2 itself computes the
  same way.
* On reference fixtures (SZ and SU(2)) handed back to block2 the result
  equals block2's own dumped H*psi: the handover loses nothing."""
import os
import tempfile

import numpy as np
import pytest

from b2g import effham, synth
from b2g.problem import Problem
import b2g_oracle as O
import ref_problem as R

from conftest import fixture_path

pytestmark = pytest.mark.skipif(not os.path.exists(R.B2REF),
                                reason="reference harness not built")


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("args", [(12, 150, 60, 3), (16, 400, 160, 7)])
def test_block2_on_synthetic_site(args):
    n_orb, m, dmax, seed = args
    pb = synth.middle_site(n_orb=n_orb, m=m, max_dim=dmax, seed=seed)
    rng = np.random.default_rng(seed)
    xs = rng.standard_normal((2, pb.ket_len))
    orc = O.Oracle(pb)
    with tempfile.TemporaryDirectory() as d:
        R.write_problem_dir(pb, d, xs)
        out = R.run_reference(d, pb.sym, threads=8, reps=1)
    hx = out["hx"].reshape(2, -1)
    for r in range(2):
        assert rel(hx[r], orc.hpsi(xs[r], nthreads=8)) < 1e-12, r
    assert rel(out["diag"], orc.diag()) < 1e-13
    h = effham.host_plan(pb)
    try:
        assert out["nflop_mac"] == effham.host_plan_info(h)["macs"]
    finally:
        effham.destroy(h)


@pytest.mark.parametrize("name", ["hub8_sz", "n2_su2"])
def test_block2_on_handed_back_fixture(name):
    pb = Problem.from_npz(fixture_path(name))
    g = pb.golden
    xs = np.stack([g["psi"], g["x2"]])
    with tempfile.TemporaryDirectory() as d:
        R.write_problem_dir(pb, d, xs)
        out = R.run_reference(d, pb.sym, threads=4, reps=1)
    hx = out["hx"].reshape(2, -1)
    assert rel(hx[0], g["hpsi"]) < 1e-13
    assert rel(hx[1], g["hx2"]) < 1e-13
    assert rel(out["diag"], g["diag"]) < 1e-13
    assert out["nflop_mac"] == int(pb.meta["nflop_mac"])
