import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "block2-preview_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))
GOLDEN = os.path.join(REPO, "tests", "golden")
FIXTURES = ["hub8_sz", "hub8_su2", "n2_su2", "n2_sz", "hub16_sz_m200"]
# effective Hamiltonians with a delayed operator tensor (three-tensor
# strings; reference delayed contraction), left and right
DELAYED_FIXTURES = ["n2_su2_dleft", "n2_su2_dright", "hub8_sz_dleft",
                    "hub8_sz_dright"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def fixture_path(name):
    return os.path.join(GOLDEN, name + ".npz")


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    from b2g.effham import Context
    ctx = Context(0)
    yield ctx
    ctx.close()
