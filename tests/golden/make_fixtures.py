"""Regenerate the golden fixtures under tests/golden/ (this container only).

Runs the reference harness oracle/_ref/b2ref (built by `make -C oracle ref`
from the reference's own sources) on small systems and packs each dumped
effective-Hamiltonian problem into a compressed .npz:

  <name>.npz   ket/bra infos, left/right operator infos and data, the H_eff
               strings, the reference ConnectionInfo lists (wfn, diag), psi,
               the reference H*psi of psi and of a seeded random vector, the
               diagonal, a reference Davidson energy (+ vector) and metadata.
  energies.json  reference DMRG energies of the runs (known answers).

Inputs: FCIDUMPs shipped in the reference's data/ directory, plus the 1D
Hubbard chain L=16, U=4, t=1 written here (the shipped HUBBARD-L16 has U=2).
"""
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
B2REF = os.path.join(REPO, "oracle", "_ref", "b2ref")


def write_hubbard(path, n, u, t):
    with open(path, "w") as f:
        f.write(" &FCI NORB=%3d,NELEC=%3d,MS2=  0,\n  ORBSYM=%s\n  ISYM=1,\n /\n"
                % (n, n, "1," * n))
        for i in range(1, n + 1):
            f.write(" %22.16f %3d %3d %3d %3d\n" % (u, i, i, i, i))
        for i in range(1, n):
            f.write(" %22.16f %3d %3d %3d %3d\n" % (-t, i + 1, i, 0, 0))
        f.write(" %22.16f %3d %3d %3d %3d\n" % (0.0, 0, 0, 0, 0))


def abinitio_integrals(n, seed):
    """The seeded synthetic ab initio Hamiltonian: orbital energies e_i
    sorted uniform in [-2, 1]; h = diag(e) + symmetric Gaussian
    off-diagonals (sigma 0.05); (ij|kl) = sum_P B^P_ij B^P_kl over 3n
    symmetric Gaussian B^P scaled so that E[(ii|ii)] = 0.5 (eight-fold
    symmetric, positive semi-definite).  numpy default_rng(seed), drawn in
    this order: e, the off-diagonals, the B^P."""
    rng = np.random.default_rng(seed)
    e = np.sort(rng.uniform(-2.0, 1.0, n))
    g = np.triu(rng.normal(0.0, 0.05, (n, n)), 1)
    h = np.diag(e) + g + g.T
    npp = 3 * n
    # B = (X + X^T) / 2 with X ~ N(0, s^2): diagonal variance s^2, so
    # E[(ii|ii)] = 3n s^2 = 0.5
    sd = np.sqrt(0.5 / npp)
    x = rng.normal(0.0, sd, (npp, n, n))
    b = 0.5 * (x + x.transpose(0, 2, 1))
    eri = np.einsum("pij,pkl->ijkl", b, b)
    return h, eri


def write_abinitio(path, n, nelec, seed):
    """FCIDUMP of abinitio_integrals(n, seed): singlet, no point group, the
    unique integrals of the eight-fold symmetry (i >= j, k >= l, ij >= kl)
    and the one-electron ones (i >= j), core energy 0."""
    h, eri = abinitio_integrals(n, seed)
    with open(path, "w") as f:
        f.write(" &FCI NORB=%3d,NELEC=%3d,MS2=  0,\n  ORBSYM=%s\n  ISYM=1,\n"
                " /\n" % (n, nelec, "1," * n))
        for i in range(n):
            for j in range(i + 1):
                for k in range(n):
                    for l in range(k + 1):
                        if i * (i + 1) // 2 + j < k * (k + 1) // 2 + l:
                            continue
                        f.write(" %22.16e %3d %3d %3d %3d\n" % (
                            eri[i, j, k, l], i + 1, j + 1, k + 1, l + 1))
        for i in range(n):
            for j in range(i + 1):
                f.write(" %22.16e %3d %3d %3d %3d\n" % (h[i, j], i + 1,
                                                        j + 1, 0, 0))
        f.write(" %22.16e %3d %3d %3d %3d\n" % (0.0, 0, 0, 0, 0))


def run(name, fcidump, sym, bdims, noises, nsweeps, dump_sweep, dump_site,
        work, dav_conv="1e-9", delayed=False, names=None):
    """One reference DMRG run; dumps site `dump_site` (an int, or a list of
    sites written to the fixtures `names`).  delayed=True turns on the
    reference's delayed contraction of the environments
    (MovingEnvironment::delayed_contraction = normal_ops, the Python
    driver's default), whose effective Hamiltonians hold three-tensor
    strings."""
    out = os.path.join(work, name)
    if os.path.exists(out):
        shutil.rmtree(out)
    sites = dump_site if isinstance(dump_site, list) else [dump_site]
    names = names or [name]
    cmd = [B2REF, "--fcidump", fcidump, "--sym", sym, "--out", out,
           "--bdims", ",".join(map(str, bdims)),
           "--noises", ",".join(map(str, noises)),
           "--nsweeps", str(nsweeps), "--threads", "8",
           "--dump-sweep", str(dump_sweep),
           "--dump-sites", ",".join(map(str, sites)),
           "--dav-conv", dav_conv, "--delayed", "1" if delayed else "0"]
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    energies = {}
    with open(os.path.join(out, "energies.txt")) as f:
        for line in f:
            p = line.split()
            if p[0] == "final":
                energies["final"] = float(p[1])
    for site, fname in zip(sites, names):
        pack(os.path.join(out, "site%d_sw%d" % (site, dump_sweep)), fname)
    return energies


def pack(site_dir, name):
    if os.path.isdir(site_dir):
        arrs = {}
        for fn in sorted(os.listdir(site_dir)):
            if fn.endswith(".npy"):
                arrs[fn[:-4]] = np.load(os.path.join(site_dir, fn))
            elif fn.endswith(".smat"):
                # block2 sparse-matrix files, kept byte for byte
                with open(os.path.join(site_dir, fn), "rb") as f:
                    arrs[fn[:-5] + "_smat"] = np.frombuffer(f.read(),
                                                            np.uint8)
        meta = {}
        with open(os.path.join(site_dir, "meta.txt")) as f:
            for line in f:
                k, v = line.split()
                meta[k] = v
        arrs["meta_json"] = np.frombuffer(json.dumps(meta).encode(), np.uint8)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrs)
        print("  wrote", name + ".npz", flush=True)


def delayed_fixtures(work):
    """Sites whose effective Hamiltonian has a delayed left (site 3) or
    delayed right (site 4) operator tensor, N2/STO-3G SU(2) and Hubbard
    L=8 SZ."""
    data = os.path.join(REF, "data")
    res = {}
    res["n2_su2_delayed"] = run(
        "n2_su2_delayed", os.path.join(data, "N2.STO3G.FCIDUMP"), "su2",
        [100], [1e-6, 1e-7, 0], 8, 2, [3, 4], work, delayed=True,
        names=["n2_su2_dleft", "n2_su2_dright"])
    res["hub8_sz_delayed"] = run(
        "hub8_sz_delayed", os.path.join(data, "HUBBARD-L8.FCIDUMP"), "sz",
        [200], [1e-6, 0], 8, 1, [3, 4], work, delayed=True,
        names=["hub8_sz_dleft", "hub8_sz_dright"])
    return res


def main():
    # raw harness dumps go to a scratch directory removed afterwards; only
    # the packed fixtures are kept
    with tempfile.TemporaryDirectory(prefix="b2g_fixtures_") as work:
        return build_all(work)


def build_all(work):
    if "--delayed-only" in sys.argv:
        res = delayed_fixtures(work)
        with open(os.path.join(HERE, "energies.json")) as f:
            old = json.load(f)
        old.update(res)
        with open(os.path.join(HERE, "energies.json"), "w") as f:
            json.dump(old, f, indent=1, sort_keys=True)
        print(json.dumps(res, indent=1))
        return 0
    hub16 = os.path.join(work, "HUBBARD-L16-U4.FCIDUMP")
    write_hubbard(hub16, 16, 4.0, 1.0)
    data = os.path.join(REF, "data")
    res = {}
    res["hub8_sz"] = run("hub8_sz", os.path.join(data, "HUBBARD-L8.FCIDUMP"),
                         "sz", [200], [1e-6, 0], 8, 1, 3, work)
    res["hub8_su2"] = run("hub8_su2", os.path.join(data, "HUBBARD-L8.FCIDUMP"),
                          "su2", [200], [1e-6, 0], 8, 1, 3, work)
    res["n2_su2"] = run("n2_su2", os.path.join(data, "N2.STO3G.FCIDUMP"),
                        "su2", [100], [1e-6, 1e-7, 0], 8, 2, 4, work)
    res["n2_sz"] = run("n2_sz", os.path.join(data, "N2.STO3G.FCIDUMP"),
                       "sz", [60], [1e-6, 1e-7, 0], 8, 2, 4, work)
    res["hub16_sz_m200"] = run("hub16_sz_m200", hub16, "sz", [200],
                               [1e-5, 1e-5, 1e-5, 1e-5, 0], 10, 4, 7, work,
                               dav_conv="1e-8")
    # energy-only runs (known answers)
    res.update(delayed_fixtures(work))
    res["n2_su2_m500"] = run("n2_su2_m500",
                             os.path.join(data, "N2.STO3G.FCIDUMP"), "su2",
                             [500], [1e-6, 1e-7, 0], 10, -1, -1, work)
    with open(os.path.join(HERE, "energies.json"), "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    sys.exit(main())
