"""The block2 binding of libb2g.so (block2-preview_amd/integration/
b2g_block2.hpp), CPU only, in this container where the reference sources
are present:

* it compiles against block2's headers (src/block2_core.hpp,
  src/block2_dmrg.hpp) and include/b2g.h for SZ and SU(2);
* run inside the reference's own DMRG (oracle/_ref/b2shim, the reference
  harness with the binding compiled in), the b2g_problem it marshals from
  the live EffectiveHamiltonian, built host-only by b2g_effham_create,
  gives the reference's ConnectionInfo (src/core/sparse_matrix.hpp:163-292)
  entry by entry and the reference's multiply-add count, for plain
  OpProduct strings and for the delayed three-tensor OpSumProd strings of
  the reference's delayed contraction (left and right)."""
import os
import subprocess
import tempfile

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
SHIM_DIR = os.path.join(REPO, "block2-preview_amd", "integration")
B2SHIM = os.path.join(REPO, "oracle", "_ref", "b2shim")

needs_ref = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")),
                               reason="reference sources not present")

TU = r"""
#include "block2_core.hpp"
#include "block2_dmrg.hpp"
#include "b2g_block2.hpp"
using namespace block2;
template struct b2g_block2::ProblemPack<SZ>;
template struct b2g_block2::ProblemPack<SU2>;
template std::tuple<double, int, size_t, double>
b2g_block2::eigs<SZ>(EffectiveHamiltonian<SZ, double> &, double, int, int,
                     int, int);
template std::tuple<double, int, size_t, double>
b2g_block2::eigs<SU2>(EffectiveHamiltonian<SU2, double> &, double, int, int,
                      int, int);
"""


@needs_ref
def test_binding_compiles_against_block2():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "tu.cpp")
        with open(src, "w") as f:
            f.write(TU)
        r = subprocess.run(
            ["g++", "-std=c++17", "-fsyntax-only", "-fopenmp", "-w",
             "-D_USE_CORE", "-D_USE_DMRG", "-D_USE_SU2SZ", "-D_LARGE_BOND",
             "-D_F77UNDERSCORE", "-I" + os.path.join(REF, "src"),
             "-I" + os.path.join(REPO, "include"), "-I" + SHIM_DIR, src],
            capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]


def _b2shim():
    if not os.path.exists(B2SHIM):
        subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"),
                        "shimcheck"], check=True)
    return B2SHIM


CASES = [
    # (FCIDUMP in the reference's data/, symmetry, M, dumped sites, delayed)
    ("HUBBARD-L8.FCIDUMP", "sz", 50, [2, 4], 0),
    ("HUBBARD-L8.FCIDUMP", "sz", 50, [2, 4], 1),
    ("N2.STO3G.FCIDUMP", "su2", 60, [3], 0),
    ("N2.STO3G.FCIDUMP", "su2", 60, [3, 6], 1),
]


@needs_ref
@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s-d%d" % (
    c[0].split(".")[0], c[1], c[4]))
def test_binding_marshals_reference_problems(case):
    fcidump, sym, m, sites, delayed = case
    exe = _b2shim()
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "run")
        cmd = [exe, "--fcidump", os.path.join(REF, "data", fcidump),
               "--sym", sym, "--out", out, "--bdims", str(m), "--noises", "0",
               "--nsweeps", "2", "--threads", "4", "--dump-sweep", "1",
               "--dump-sites", ",".join(map(str, sites)), "--dav-conv",
               "1e-9", "--delayed", str(delayed)]
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL, timeout=600)
        n_sumprod = 0
        for s in sites:
            with open(os.path.join(out, "site%d_sw1" % s, "shim.txt")) as f:
                nt, nsp, nc, eq, fdiff, macs, ref_macs = f.read().split()
            assert int(nt) > 0
            assert int(nc) > 0 and int(eq) == int(nc), (case, s, nc, eq)
            assert float(fdiff) < 1e-13, (case, s, fdiff)
            assert int(macs) == int(ref_macs), (case, s)
            n_sumprod += int(nsp)
        # the delayed runs carry three-tensor strings
        assert (n_sumprod > 0) == bool(delayed), (case, n_sumprod)
