"""The device H*psi of a problem read from block2's EFFPROB1 file.

Each fixture keeps the reference's own EffectiveProblem::save_problem_file
output for the dumped site; the Problem b2g.effprob reads from it, with the
dumped operator data attached (matched to its operator list through the H
strings that use them), gives
the reference's H*psi of both dumped vectors (FP64, 1e-12 relative) and
its diagonal (1e-12 relative), through the same entry points as any other
problem.
"""
import numpy as np
import pytest
import torch

from conftest import DELAYED_FIXTURES, FIXTURES, fixture_path

from b2g import effprob as EP
from b2g.effham import EffectiveHamiltonian
from b2g.problem import Problem
from test_effprob import attach

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def dev(x):
    return torch.tensor(np.ascontiguousarray(x), dtype=torch.float64,
                        device="cuda:0")


@pytest.mark.parametrize("name", FIXTURES + DELAYED_FIXTURES)
def test_effprob_hpsi_matches_reference(gpu_ctx, name):
    pb = Problem.from_npz(fixture_path(name))
    g = pb.golden
    if "eff_problem_smat" not in g:
        pytest.skip("fixture without an EFFPROB1 file")
    print("case", name, flush=True)
    fp = EP.to_problem(EP.EffProblemFile(bytes(g["eff_problem_smat"])))
    attach(fp, pb)
    h = EffectiveHamiltonian(fp, gpu_ctx)
    try:
        assert rel(h.multiply(dev(g["psi"])).cpu().numpy(), g["hpsi"]) < 1e-12
        assert rel(h.multiply(dev(g["x2"])).cpu().numpy(), g["hx2"]) < 1e-12
        assert rel(h.diag.cpu().numpy(), g["diag"]) < 1e-12
    finally:
        h.deallocate()


@pytest.mark.parametrize("name", FIXTURES + DELAYED_FIXTURES)
def test_saved_files_hpsi_matches_reference(gpu_ctx, name):
    """The problem assembled from block2's files alone: EFFPROB1 structure
    plus the environments' OperatorTensor files (and the delayed side's
    sub-operator tensors)."""
    pb = Problem.from_npz(fixture_path(name))
    g = pb.golden
    if "eff_problem_smat" not in g or "lopt_smat" not in g:
        pytest.skip("fixture without the saved files")
    print("case", name, flush=True)
    fp = EP.to_problem(EP.EffProblemFile(bytes(g["eff_problem_smat"])))
    ts = [EP.tensor_from_bytes(bytes(g[k])) if k in g else None
          for k in ("lopt_smat", "ropt_smat", "dlopt_smat", "dropt_smat")]
    EP.attach_operator_data(fp, *ts)
    h = EffectiveHamiltonian(fp, gpu_ctx)
    try:
        assert rel(h.multiply(dev(g["psi"])).cpu().numpy(), g["hpsi"]) < 1e-12
        assert rel(h.multiply(dev(g["x2"])).cpu().numpy(), g["hx2"]) < 1e-12
    finally:
        h.deallocate()
