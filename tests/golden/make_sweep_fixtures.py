"""Sweep fixtures: block2's own sweep inputs and results.

Runs the reference harness oracle/_ref/b2ref (built by `make -C oracle
ref`; needs the reference sources, so it runs in the build container only)
and copies into tests/golden/sweep_<case>/:

  mpo.bin.gz, mps_info.bin, mps_<i>.bin, mps.txt   the MPO (gzip) and the
      starting MPS in block2's file formats (MPO::save_data, MPSInfo::save_data,
      SparseMatrix::save_data with info), written before the first sweep
  steps.npy     per sweep step (7 columns): sweep, step, Davidson energy,
                discarded weight, the generator's next 32-bit output, the
                new bond basis's total states and a checksum of its
                (quantum, count) pairs (block2's DMRG::sweep, a run without
                dumps so that the noise stream is block2's own)
  settings.txt  bond dimensions, noises, Davidson thresholds, tol, seeds
  env_s<sweep>_<site>/   from separate runs with dumps at those sites:
      lenv.bin / renv.bin   the environments the effective Hamiltonian was
                            blocked from (OperatorTensor::save_data)
      lblk.bin / rblk.bin   the blocked operators of the effective
                            Hamiltonian
      mps_left.bin / mps_right.bin   the MPS tensors left of the center and
                            right of the two sites
      split.npz             block2's noisy truncations of the site's
                            Davidson vector (harness cases sW, sD, sE, sF:
                            density matrix, factors, error, settings) with
                            the vector and its layout

The random initial MPS is seeded (--seed) and the generator re-seeded
before the sweeps (--noise-seed), so block2's density-matrix noise can be
reproduced.  Inputs: FCIDUMPs shipped in the reference's data/ directory.

    python tests/golden/make_sweep_fixtures.py
"""
import gzip
import os
import shutil

import numpy as np
import subprocess
import sys
import tempfile

from make_fixtures import write_abinitio, write_hubbard

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
B2REF = os.path.join(REPO, "oracle", "_ref", "b2ref")

CASES = {
    # name: fcidump, sym, bdims, noises, sweeps, dump runs (sweep, sites)
    "hub8_sz": ("HUBBARD-L8.FCIDUMP", "sz", [100], [1e-6, 0], 6,
                [(0, [0, 1, 2, 3]), (1, [6, 5, 4])]),
    "hub8_su2": ("HUBBARD-L8.FCIDUMP", "su2", [100], [1e-6, 0], 6,
                 [(0, [0, 1, 2, 3]), (1, [6, 5, 4])]),
    # the same sweeps without noise: every step is then fixed by the
    # Davidson threshold alone (with noise, the noise is drawn in a basis
    # whose eigenvector signs are the eigensolver's own choice)
    "hub8_sz_quiet": ("HUBBARD-L8.FCIDUMP", "sz", [100], [0], 4, []),
    "hub8_su2_quiet": ("HUBBARD-L8.FCIDUMP", "su2", [100], [0], 4, []),
    # the L=16 chain at U=4, bond dimension 200, noise 1e-5 for four sweeps
    # then none, ten sweeps (the FCIDUMP is written here: the shipped
    # HUBBARD-L16 has U=2)
    "hub16_sz": ("HUBBARD-L16-U4.FCIDUMP", "sz", [200],
                 [1e-5, 1e-5, 1e-5, 1e-5, 0], 10, []),
    # the same with the Python driver's default noise
    # (ReducedPerturbativeCollected)
    "hub16_sz_pert": ("HUBBARD-L16-U4.FCIDUMP", "sz", [200],
                      [1e-5, 1e-5, 1e-5, 1e-5, 0], 10, [], 0.0, "pert"),
    # the seeded synthetic ab initio Hamiltonian at n=26 orbitals, 26
    # electrons, singlet (make_fixtures.write_abinitio, seed 2026), without
    # noise: bond dimension 50, then 100, then 250
    "c2_su2": ("C2-N26-SEED2026.FCIDUMP", "su2", [50, 100, 250], [0], 3,
               []),
    # perturbative noise (ReducedPerturbativeCollected, the Python
    # driver's default): no random numbers, so noisy steps are comparable
    "hub8_sz_pert": ("HUBBARD-L8.FCIDUMP", "sz", [100], [1e-5, 0], 4, [],
                     0.0, "pert"),
    "hub8_su2_pert": ("HUBBARD-L8.FCIDUMP", "su2", [100], [1e-5, 0], 4, [],
                      0.0, "pert"),
    "n2_su2_pert": ("N2.STO3G.FCIDUMP", "su2", [60], [1e-5, 0], 4, [],
                    0.0, "pert"),
    "n2_su2": ("N2.STO3G.FCIDUMP", "su2", [60], [1e-6, 0], 6,
               [(0, [0, 1, 2, 3, 4, 5]), (1, [8, 7, 6, 5, 4, 3])]),
}
FILES = ["lenv.bin", "renv.bin", "lblk.bin", "rblk.bin", "mps_left.bin",
         "mps_right.bin"]


def run(fcidump, sym, bdims, noises, nsweeps, out, dump=None, thrd=0.0,
        noise_type="dm"):
    cmd = [B2REF, "--fcidump", fcidump, "--sym", sym, "--out", out,
           "--bdims", ",".join(map(str, bdims)),
           "--noises", ",".join(map(str, noises)),
           "--nsweeps", str(nsweeps), "--threads", "8", "--dav-conv", "1e-8",
           "--save-sweep", "1", "--seed", "1234", "--noise-seed", "4321"]
    if thrd:
        cmd += ["--dav-thrd", repr(thrd)]
    if noise_type != "dm":
        cmd += ["--noise-type", noise_type]
    if dump is not None:
        cmd += ["--dump-sweep", str(dump[0]),
                "--dump-sites", ",".join(map(str, dump[1]))]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL)


def main():
    names = sys.argv[1:] or list(CASES)
    for name in names:
        fd, sym, bdims, noises, ns, dumps = CASES[name][:6]
        thrd = CASES[name][6] if len(CASES[name]) > 6 else 0.0
        ntype = CASES[name][7] if len(CASES[name]) > 7 else "dm"
        dst = os.path.join(HERE, "sweep_" + name)
        shutil.rmtree(dst, ignore_errors=True)
        os.makedirs(dst)
        with tempfile.TemporaryDirectory(prefix="b2g_sweep_") as tmp:
            if fd == "HUBBARD-L16-U4.FCIDUMP":
                fcidump = os.path.join(tmp, fd)
                write_hubbard(fcidump, 16, 4.0, 1.0)
            elif fd == "C2-N26-SEED2026.FCIDUMP":
                fcidump = os.path.join(tmp, fd)
                write_abinitio(fcidump, 26, 26, 2026)
            else:
                fcidump = os.path.join(REF, "data", fd)
            run(fcidump, sym, bdims, noises, ns, os.path.join(tmp, "clean"),
                thrd=thrd, noise_type=ntype)
            src = os.path.join(tmp, "clean", "sweep")
            for f in sorted(os.listdir(src)):
                if f == "mpo.bin":
                    # kept gzip-compressed (the readers take either form)
                    with open(os.path.join(src, f), "rb") as fi:
                        raw = fi.read()
                    with open(os.path.join(dst, f + ".gz"), "wb") as fo:
                        fo.write(gzip.compress(raw, 9, mtime=0))
                    continue
                shutil.copy(os.path.join(src, f), os.path.join(dst, f))
            for k, dump in enumerate(dumps):
                out = os.path.join(tmp, "dump%d" % k)
                run(fcidump, sym, bdims, noises, ns, out, dump, thrd, ntype)
                for site in dump[1]:
                    d = os.path.join(out, "site%d_sw%d" % (site, dump[0]))
                    e = os.path.join(dst, "env_s%d_%d" % (dump[0], site))
                    os.makedirs(e)
                    for f in FILES:
                        if os.path.exists(os.path.join(d, f)):
                            shutil.copy(os.path.join(d, f),
                                        os.path.join(e, f))
                    arrs = {}
                    for f in sorted(os.listdir(d)):
                        if f.endswith(".npy") and (
                                f.startswith("ket_") or f == "psi_dav.npy" or
                                f[:3] in ("sW_", "sD_", "sE_", "sF_")):
                            arrs[f[:-4]] = np.load(os.path.join(d, f))
                    np.savez_compressed(os.path.join(e, "split.npz"), **arrs)
        print(name, "->", dst)
    return 0


if __name__ == "__main__":
    sys.exit(main())
