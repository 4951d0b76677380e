"""The host worker threads that build the window plans (record generation,
grouping, staging copies) are created once per process.  A process forked
from one that already used them has none of those threads: its first plan
build must start workers of its own instead of waiting on the parent's
(which would never run).  Host-only blocking plans of a sweep fixture are
built in the parent, then in a forked child, and checked against the oracle
in both (CPU only), with B2G_PLAN_JOB_ITEMS set so low that even these
small plans are split over several workers."""
import multiprocessing as mp
import os

import pytest

from sweep_common import case_dir
import test_env_oracle as T

CASE, ENV = "hub8_sz", "env_s0_2"


def _build_and_check():
    from b2g import mpo as M
    mpos = {CASE: M.load_mpo(os.path.join(case_dir(CASE), "mpo.bin"))}
    T.test_host_tp_connections_match_oracle(CASE, ENV, True, mpos)
    T.test_host_tp_connections_match_oracle(CASE, ENV, False, mpos)


def _child():
    try:
        _build_and_check()
    except BaseException:
        os._exit(1)
    os._exit(0)


@pytest.mark.skipif(not os.path.isdir(os.path.join(case_dir(CASE), ENV)),
                    reason="sweep fixture absent")
def test_plans_in_forked_child_use_workers_of_their_own(monkeypatch):
    # small plans split into several host jobs, so that the workers start
    monkeypatch.setenv("B2G_PLAN_JOB_ITEMS", "16")
    _build_and_check()                      # the parent's workers exist now
    p = mp.get_context("fork").Process(target=_child)
    p.start()
    p.join(120)
    if p.is_alive():
        p.kill()
        p.join()
        pytest.fail("plan build in the forked child did not finish")
    assert p.exitcode == 0
