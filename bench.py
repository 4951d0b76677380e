"""Benchmark of the two-site DMRG effective-Hamiltonian hot path on one GPU.

Workload: the batched H*psi microbenchmark -- one middle-site effective
Hamiltonian of the n=40-orbital, M=3000, SU(2) singlet problem with its
sector-size profile reproduced from a seed (about 170 wavefunction blocks,
block dimensions 1..600, heavy tailed), 4 right-hand sides (state-averaged
roots) per H*psi.  A step is one H*psi over the 4 vectors (FP64 MFMA),
with every input already resident in HBM.

Printed JSON (one line): whole-job FP64 H*psi throughput in TFLOP/s
(FLOPs = 2 * the reference's multiply-add count of the recorded pair list,
src/core/batch_gemm.hpp:307), the roofline of the dominant kernel measured
with HIP events on the library's stream, the FP32 throughput, a Davidson
solve in FP64 and in mixed precision (FP32 matvecs + FP64 refinement) with
|dE| and 1-|<psi_mixed|psi_fp64>|, and the CPU baseline: block2 itself
(oracle/_ref/b2ref, built from the reference's sources) timed on one full
H*psi of the same site on the host cores, with its result compared to the
GPU's, beside the oracle's OpenMP restatement on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "block2-preview_amd"))

METRIC = ("H·ψ TFLOP/s + s/sweep; FP32-mixed vs FP64 GPU at "
          "|ΔE|≤1e-6 Eh")
FP64_MFMA_PEAK = 78.6   # TFLOP/s, MI355X dense FP64 matrix
FP32_MFMA_PEAK = 157.3  # TFLOP/s, MI355X dense FP32 matrix


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE + WRITE_SIZE summary (profiles/*.md, copy-calibrated), or
    None when no profile of that kernel is committed."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(REPO, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                m = re.match(r"traffic_bytes_per_launch (\S+) (\S+)", line)
                if m and m.group(1) == kernel:
                    return float(m.group(2))
    return None


def cpu_baseline(pb, ldata, rdata, seconds, threads):
    """Oracle restatement timed on a bounded prefix of the recorded pairs:
    a 64-pair calibration run gives the cost per pair, then a prefix sized
    for about `seconds` of CPU work is timed."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import numpy as np
    import b2g_oracle
    pb.lops.data = ldata
    pb.rops.data = rdata
    orc = b2g_oracle.Oracle(pb)
    x = np.random.default_rng(11).standard_normal(pb.ket_len)
    t0 = time.perf_counter()
    _, m0 = orc.hpsi_sample(x, 64, threads)
    dt0 = max(time.perf_counter() - t0, 1e-3)
    n = int(max(64, min(seconds * (m0 / dt0) / max(m0 / 64.0, 1.0), 1 << 30)))
    t0 = time.perf_counter()
    _, macs = orc.hpsi_sample(x, n, threads)
    dt = time.perf_counter() - t0
    return dict(value=round(2.0 * macs / dt / 1e12, 6), unit="TFLOP/s",
                cores=threads, kind="port",
                sample="first %d recorded pairs of one H*psi (1 RHS, %.3g of "
                       "%.3g multiply-adds), %.1f s" % (
                           n, macs, orc.macs, dt))


def cpu_reference(pb, ldata, rdata, y_gpu_fn, threads):
    """block2 itself (oracle/_ref/b2ref, built from the reference's sources,
    MKL DGEMM, `threads` OpenMP threads) on this site: the site is handed
    over as arrays, block2 builds its own EffectiveHamiltonian and runs its
    Tasked replay (the path of EffectiveHamiltonian::eigs) on one seeded
    vector, timed twice (best taken).  Also returns the relative difference
    between block2's H x and the GPU's FP64 H x of the same vector."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import shutil
    import tempfile
    import numpy as np
    import ref_problem
    if not os.path.exists(ref_problem.B2REF):
        return None
    x = np.random.default_rng(11).standard_normal(pb.ket_len)
    d = tempfile.mkdtemp(prefix="b2g_ref_")
    try:
        ref_problem.write_problem_dir(pb, d, x[None], ldata, rdata)
        out = ref_problem.run_reference(d, pb.sym, threads, reps=2)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    t = min(out["t_hpsi_s"])
    y = y_gpu_fn(x)
    hx = out["hx"]
    rel = float(np.linalg.norm(hx - y) / np.linalg.norm(hx))
    return dict(value=round(2.0 * out["nflop_mac"] / t / 1e12, 6),
                unit="TFLOP/s", cores=threads, kind="reference",
                sample="one full H*psi of this site (1 RHS, %d multiply-adds) "
                       "by block2's Tasked replay with MKL DGEMM, best of 2: "
                       "%.3f s; effective-Hamiltonian construction %.2f s" % (
                           out["nflop_mac"], t, out["t_ctor_s"]),
                gpu_fp64_vs_block2_rel=rel)


C1_DIR = os.path.join(REPO, "tests", "golden", "sweep_hub16_sz_pert")


def _hubbard_fcidump(path, n, u, t):
    """FCIDUMP of the open 1D Hubbard chain (hopping -t, on-site U)."""
    with open(path, "w") as f:
        f.write(" &FCI NORB=%3d,NELEC=%3d,MS2=  0,\n  ORBSYM=%s\n  ISYM=1,\n"
                " /\n" % (n, n, "1," * n))
        for i in range(n):
            f.write(" %22.16f %3d %3d %3d %3d\n" % (u, i + 1, i + 1, i + 1,
                                                    i + 1))
        for i in range(n - 1):
            f.write(" %22.16f %3d %3d %3d %3d\n" % (-t, i + 2, i + 1, 0, 0))
        f.write(" %22.16f %3d %3d %3d %3d\n" % (0.0, 0, 0, 0, 0))


def sweep_leg(ctx, threads, run_cpu):
    """Whole two-site sweeps of the L=16, U=4 Hubbard chain (SZ, M=200,
    noise 1e-5 for four sweeps then none, block2's default
    ReducedPerturbativeCollected noise, Davidson thresholds as block2
    sets them) from block2's own MPO and starting MPS, through the sweep
    driver on the device: per-sweep seconds and the final energy against
    block2's.  Beside it, block2 itself (oracle/_ref/b2ref, MKL, `threads`
    OpenMP threads) runs the same sweeps from the same seeded MPS on the
    host."""
    import numpy as np
    from b2g import dmrg as D
    st = {}
    with open(os.path.join(C1_DIR, "settings.txt")) as f:
        for line in f:
            p = line.split()
            st[p[0]] = p[1:]
    args = (int(st["nsweeps"][0]), True, float(st["tol"][0]),
            [float(x) for x in st["dav_conv"]])

    def make():
        return D.DMRG(ctx, "sz", os.path.join(C1_DIR, "mpo.bin"), C1_DIR,
                      [int(x) for x in st["bdims"]],
                      [float(x) for x in st["noises"]],
                      noise_type=st["noise_type"][0],
                      noise_seed=int(st["noise_seed"][0]))
    # a first solve loads the code objects and grows the caching allocator;
    # the second, from the same starting MPS, is the measured one
    make().solve(*args)
    dm = make()
    t0 = time.perf_counter()
    e = dm.solve(*args)
    wall = time.perf_counter() - t0
    ref = float(st["final"][0])
    steps = np.array([s[5] for s in dm.steps])
    out = dict(workload="two-site DMRG sweeps, 1D Hubbard L=16, U=4, t=1, "
                        "half filling, SZ, M=200, perturbative noise 1e-5 "
                        "x4 then 0, block2's MPO and starting MPS",
               sweeps=len(dm.sweep_time),
               s_per_sweep=[round(t, 4) for t in dm.sweep_time],
               mean_s_per_sweep=round(float(np.mean(dm.sweep_time)), 4),
               total_s=round(wall, 4), final_energy=e,
               block2_final_energy=ref, abs_dE_vs_block2=abs(e - ref),
               step_s_mean=round(float(steps.mean()), 5),
               step_s_max=round(float(steps.max()), 5),
               phase_s={k: round(v, 4) for k, v in dm.phase_time.items()})
    log("sweeps (L=16 Hubbard, M=200): %d sweeps, %.3f s/sweep mean, "
        "E %.12f (block2 %.12f)" % (out["sweeps"], out["mean_s_per_sweep"],
                                     e, ref))
    if run_cpu:
        import shutil
        import subprocess
        import tempfile
        b2ref = os.path.join(REPO, "oracle", "_ref", "b2ref")
        if os.path.exists(b2ref):
            d = tempfile.mkdtemp(prefix="b2g_c1_")
            try:
                fd = os.path.join(d, "H16.FCIDUMP")
                _hubbard_fcidump(fd, 16, 4.0, 1.0)
                cmd = [b2ref, "--fcidump", fd, "--sym", "sz", "--out",
                       os.path.join(d, "out"), "--bdims", st["bdims"][0],
                       "--noises", ",".join(st["noises"]), "--nsweeps",
                       st["nsweeps"][0], "--threads", str(threads),
                       "--dav-conv", "1e-8", "--seed", "1234",
                       "--noise-seed", st["noise_seed"][0],
                       "--noise-type", st["noise_type"][0]]
                t0 = time.perf_counter()
                subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL,
                               stderr=subprocess.DEVNULL, timeout=600)
                cwall = time.perf_counter() - t0
                en = {}
                with open(os.path.join(d, "out", "energies.txt")) as f:
                    for line in f:
                        p = line.split()
                        en[p[0]] = p[1:]
                ss = [float(x) for x in en.get("sweep_s", [])]
                out["block2_cpu"] = dict(
                    cores=threads, sweeps=len(ss),
                    s_per_sweep=[round(t, 4) for t in ss],
                    mean_s_per_sweep=round(float(np.mean(ss)), 4) if ss
                    else None, process_s=round(cwall, 3),
                    final_energy=float(en["final"][0]))
                if ss:
                    out["speedup_vs_block2_cpu"] = round(
                        float(np.sum(ss)) / float(np.sum(dm.sweep_time)), 3)
                log("block2 on %d host cores: %s s/sweep" % (
                    threads, out["block2_cpu"]["mean_s_per_sweep"]))
            finally:
                shutil.rmtree(d, ignore_errors=True)
    return out


def davidson_profile(args):
    """One FP64 Davidson solve (one root, the bench's start vector) on the
    bench site, for a kernel-trace profile of the bandwidth-bound kernels:
    prints the algorithmic HBM bytes of each Davidson step kernel summed
    over the solve (step t of the solve works on a subspace of m = t
    vectors of |psi| doubles; the last step, converged, appends nothing)
    and of the diagonal (its output written once, the operator diagonals
    read once), to be divided by the summed kernel durations the profiler
    reports."""
    import ctypes as C
    import numpy as np
    import torch
    from b2g import _lib as L
    from b2g import synth
    from b2g.effham import Context, EffectiveHamiltonian
    pb = synth.middle_site(n_orb=args.n_orb, m=args.m, max_dim=args.max_dim,
                           seed=args.seed, pair_scale=args.pair_scale,
                           data="device")
    ctx = Context(0)
    ldev, rdev = synth.fill_device(pb, ctx.device)
    torch.cuda.synchronize()
    h = EffectiveHamiltonian(pb, ctx, fp32=False, device_data=(ldev, rdev))
    n = h.ket_len
    nt, we, re_ = C.c_uint64(), C.c_uint64(), C.c_uint64()
    L.check(L.lib().b2g_effham_diag_stats(h.h, C.byref(nt), C.byref(we),
                                          C.byref(re_)))
    # the bench's guess: the lowest diagonal state plus seeded noise
    g = torch.Generator(device=ctx.device).manual_seed(args.seed)
    x = 0.3 * torch.randn(n, generator=g, dtype=torch.float64,
                          device=ctx.device) / float(np.sqrt(n))
    x[int(torch.argmin(h.diag))] += 1.0
    x /= torch.linalg.norm(x)
    e, nd, nflop, tdav = h.eigs(x, conv_thrd=1e-9, precision="f64")
    steps = list(range(1, nd + 1))
    grow = steps[:-1]            # steps that append a vector
    b = 8.0
    per = {
        "put_sigma_kernel": b * 2 * n * len(grow),
        "gram_mfma_kernel": sum(b * 2 * m * n for m in steps),
        "rotate_mfma_kernel": sum(b * 4 * m * n for m in steps),
        "residual_kernel": b * 3 * n * len(steps),
        "olsen1_kernel": b * 5 * n * len(steps),
        "olsen2_kernel": b * 3 * n * len(steps),
        "gs_dots_kernel": sum(2 * b * (m + 1) * n for m in grow),
        "gs_update_kernel": sum(2 * b * (m + 2) * n for m in grow),
        "norm_kernel": b * n * len(grow),
        "append_kernel": b * 3 * n * len(grow),
        "diag_kernel": b * (n + re_.value),
    }
    out = {"workload": "one FP64 Davidson solve (1 root, conv 1e-9) on the "
                       "n=%d, M=%d SU(2) middle site, seed %d" % (
                           args.n_orb, args.m, args.seed),
           "psi_len": n, "davidson_steps": nd, "energy": e,
           "diag_terms": nt.value, "diag_window_elems": we.value,
           "diag_flops": 2 * we.value, "algorithmic_bytes_total": per}
    print(json.dumps(out), flush=True)
    h.deallocate()
    ctx.close()
    return 0


C2_DIR = os.path.join(REPO, "tests", "golden", "sweep_c2_su2")


def abinitio_run(prec, bond_dims):
    """One trajectory of the n=26 sweeps (see abinitio_leg) in this
    process: per-sweep seconds, energies and phases."""
    import torch
    from b2g import dmrg as D
    from b2g.effham import Context
    ctx = Context(0)
    dm = D.DMRG(ctx, "su2", os.path.join(C2_DIR, "mpo.bin"), C2_DIR,
                list(bond_dims), [0.0], precision=prec)
    t0 = time.perf_counter()
    dm.solve(len(bond_dims), True, 0.0, None)
    wall = time.perf_counter() - t0
    dav = dm.phase_time.get("davidson", 0.0)
    res = dict(
        energies=dm.energies,
        s_per_sweep=[round(t, 3) for t in dm.sweep_time],
        total_s=round(wall, 3),
        phase_s={k: round(v, 3) for k, v in dm.phase_time.items()},
        davidson_tflops=round(2.0 * dm.davidson_macs / dav / 1e12, 3)
        if dav else None)
    dm.close()
    del dm
    torch.cuda.synchronize()
    ctx.close()
    return res


def abinitio_leg(bond_dims):
    """Sweeps on the seeded n=26 ab initio Hamiltonian (SU(2), 26 electrons,
    singlet; block2's MPO and starting MPS) with the bond dimension growing
    over `bond_dims`, noise 0: per-sweep seconds, energies and the phases
    of the steps, with FP64 products and with split-BF16 products plus the
    FP64 refinement.  Each trajectory runs in a child process of its own,
    so that the second starts on a device the first has fully released."""
    import subprocess
    out = {"workload": "two-site DMRG sweeps, synthetic ab initio "
                       "Hamiltonian n=26, SU(2), bond dimension %s, no noise"
                       % "/".join(map(str, bond_dims)),
           "bond_dims": bond_dims}
    for prec in ("f64", "bf16x3"):
        r = subprocess.run(
            [sys.executable, os.path.abspath(__file__), "--abinitio-run",
             prec, "--abinitio-bdims", ",".join(map(str, bond_dims))],
            stdout=subprocess.PIPE, text=True, timeout=900)
        if r.returncode != 0:
            raise RuntimeError("n=26 sweeps (%s) failed: exit %d" % (
                prec, r.returncode))
        out[prec] = json.loads(r.stdout.strip().splitlines()[-1])
        log("n=26 sweeps %s: %s s/sweep, E %s" % (
            prec, out[prec]["s_per_sweep"],
            ", ".join("%.8f" % e for e in out[prec]["energies"])))
    out["dE_per_sweep_bf16x3_vs_fp64"] = [
        a - b for a, b in zip(out["bf16x3"]["energies"],
                              out["f64"]["energies"])]
    out["note"] = ("per-sweep energies of the two trajectories differ while "
                   "M grows (Davidson threshold 1e-10 on |r|^2, unconverged "
                   "sweeps); at a fixed effective Hamiltonian the refined "
                   "mixed solve equals the FP64 one within 1e-8 Eh "
                   "(tests/test_gpu_sweep.py)")
    return out


def winsum_profile(args):
    """n=26 sweeps (M = 50 ... 1000, FP64) with the algorithmic bytes of
    every window-sum launch (one per window-plan run with windows, in launch
    order: 8 bytes x (elements read + elements written)) printed for a
    kernel-trace profile of winsum_kernel."""
    import ctypes as C
    from b2g import _lib as L
    from b2g import dmrg as D
    from b2g.effham import Context
    ctx = Context(0)
    rec = []
    lib = L.lib()
    run = lib.b2g_winplan_run

    def counted(h, a, b, c, accumulate):
        v = [C.c_uint64() for _ in range(5)]
        L.check(lib.b2g_winplan_stats(h, *[C.byref(x) for x in v]))
        if v[0].value > 0:
            rec.append(("run", 8 * (v[2].value + v[3].value)))
        return run(h, a, b, c, accumulate)
    lib.b2g_winplan_run = counted
    dm = D.DMRG(ctx, "su2", os.path.join(C2_DIR, "mpo.bin"), C2_DIR,
                [50, 100, 250, 500, 1000], [0.0])
    dm.solve(5, True, 0.0, None)
    print(json.dumps({"workload": "n=26 SU(2) sweeps, M = 50/100/250/500/"
                                  "1000, FP64",
                      "winsum_launches": len(rec),
                      "phases": [p for p, _ in rec],
                      "algorithmic_bytes": [b for _, b in rec]}))
    dm.close()
    ctx.close()
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-orb", type=int, default=40)
    ap.add_argument("--m", type=int, default=3000)
    ap.add_argument("--max-dim", type=int, default=600)
    ap.add_argument("--pair-scale", type=int, default=4)
    ap.add_argument("--rhs", type=int, default=4)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-davidson", action="store_true")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the whole-sweep leg (L=16 Hubbard, M=200)")
    ap.add_argument("--rates", default=None,
                    help="9 comma-separated tile rates instead of autotuning "
                         "(profiling runs: keeps the tuner's launches out of "
                         "the per-kernel statistics)")
    ap.add_argument("--rates-bf16", default=None,
                    help="9 comma-separated split-BF16 tile rates instead of "
                         "autotuning them")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB "
                         "(reference kernel of known traffic for the HBM "
                         "byte counters) and exit")
    ap.add_argument("--abinitio-run", default=None, metavar="PREC",
                    help=argparse.SUPPRESS)
    ap.add_argument("--abinitio-bdims", default="50,100,250,500,1000",
                    help=argparse.SUPPRESS)
    ap.add_argument("--winsum-profile", action="store_true",
                    help="only run the n=26 sweeps in FP64 and print the "
                         "algorithmic bytes of every window-sum launch (for "
                         "a kernel-trace profile) and exit")
    ap.add_argument("--davidson-profile", action="store_true",
                    help="only run one FP64 Davidson solve on the site and "
                         "print the algorithmic bytes of its bandwidth-bound "
                         "kernels (for a kernel-trace profile) and exit")
    args = ap.parse_args()
    if args.abinitio_run:
        print(json.dumps(abinitio_run(
            args.abinitio_run,
            [int(b) for b in args.abinitio_bdims.split(",")])), flush=True)
        return 0
    if args.winsum_profile:
        return winsum_profile(args)
    if args.davidson_profile:
        return davidson_profile(args)
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written":
                          n * 8, "launches": 5}))
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU only: the hot path is a single-GPU "
              "kernel here; run with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from b2g import autotune, synth
    from b2g import effham as eh
    from b2g.effham import Context, EffectiveHamiltonian

    t0 = time.time()
    pb = synth.middle_site(n_orb=args.n_orb, m=args.m, max_dim=args.max_dim,
                           seed=args.seed, pair_scale=args.pair_scale,
                           data="device")
    ctx = Context(0)
    dev = ctx.device
    # autotune the tile shapes on the problem's own block-size histogram
    hp = eh.host_plan(pb)
    import ctypes as C
    from b2g import _lib as L
    npair = C.c_uint64()
    L.check(L.lib().b2g_effham_pairs(hp, C.byref(npair), None))
    parr = (L.b2g_pair * npair.value)()
    L.check(L.lib().b2g_effham_pairs(hp, C.byref(npair), parr))
    eh.destroy(hp)
    rep = autotune.representative_dim(np.ctypeslib.as_array(parr))
    t_tune = time.time()
    ldev, rdev = synth.fill_device(pb, dev)
    torch.cuda.synchronize()
    gx = torch.Generator(device=dev).manual_seed(args.seed + 1)
    xt = torch.randn((args.rhs, pb.ket_len), generator=gx,
                     dtype=torch.float64, device=dev)
    tuning = {}

    def tables_for(prec, fixed):
        """Rate tables to choose from for `prec`: the library's current
        one, the uniform-block measurement (FP64) and the per-shape
        measurement on this problem; the fastest product is installed."""
        cur = (C.c_double * 9)()
        L.check(L.lib().b2g_get_tile_rates_for(L.PREC[prec], cur, 9))
        tabs = {"library": list(cur)}
        if fixed is not None:
            L.check(L.lib().b2g_set_tile_rates_for(
                L.PREC[prec], (C.c_double * 9)(*fixed), 9))
            return fixed, {"given": fixed}, None
        if prec == "f64":
            tabs["uniform_blocks"] = autotune.tune(
                ctx, dim=min(max(rep, 64), 512), install=False)
        tabs["this_problem"] = autotune.tune_on_problem(
            ctx, pb, (ldev, rdev), prec, x=xt, install=False)
        names = list(tabs)
        best, times = autotune.choose(ctx, pb, (ldev, rdev),
                                      [tabs[k] for k in names], prec, x=xt)
        ms = {k: round(t, 4) for k, t in zip(names, times)}
        return tabs[names[best]], tabs, dict(chosen=names[best],
                                              ms_per_product=ms)

    fixed = ([float(r) for r in args.rates.split(",")] if args.rates
             else None)
    rates, tabs64, ch64 = tables_for("f64", fixed)
    fixed_bf = ([float(r) for r in args.rates_bf16.split(",")]
                if args.rates_bf16 else None)
    rates_bf, tabsbf, chbf = tables_for("bf16x3", fixed_bf)
    tuning = {"f64": dict(tables={k: [round(r, 2) for r in v]
                                  for k, v in tabs64.items()}, choice=ch64),
              "bf16x3": dict(tables={k: [round(r, 2) for r in v]
                                     for k, v in tabsbf.items()},
                             choice=chbf)}
    log("tiling (%.1fs): FP64 %s, split-BF16 %s" % (
        time.time() - t_tune, ch64, chbf))
    del xt
    h = EffectiveHamiltonian(pb, ctx, fp32=True, device_data=(ldev, rdev))
    log("setup %.1fs: |psi| %d, pairs %d, MACs/H*psi %.4g, terms %d" % (
        time.time() - t0, h.ket_len, h.n_pairs, h.macs, len(pb.term_conj)))
    pairs = h.pairs()
    c_m = pairs["c_m"].astype(np.float64)
    c_n = pairs["c_n"].astype(np.float64)
    v_m = pairs["v_m"].astype(np.float64)
    v_n = pairs["v_n"].astype(np.float64)
    macs1 = float(np.sum(c_m * c_n * v_n))  # work = C op(B)
    macs2 = float(np.sum(v_m * c_m * v_n))  # V += alpha op(A) work
    R = args.rhs
    g = torch.Generator(device=dev).manual_seed(args.seed)
    x = torch.randn((R, h.ket_len), generator=g, dtype=torch.float64,
                    device=dev)
    y = torch.empty_like(x)
    stream = torch.cuda.ExternalStream(ctx.lib.b2g_ctx_stream(ctx.h))

    def run(prec, steps, timed):
        for _ in range(args.warmup):
            y.copy_(h.multiply(x, precision=prec))
        torch.cuda.synchronize()
        ctx.synchronize()
        if timed:
            h.set_timing(True)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            code = h.lib.b2g_effham_hpsi(h.h, x.data_ptr(), y.data_ptr(), R,
                                         1.0, 0, L.PREC[prec])
            if code != 0:
                raise RuntimeError("b2g_effham_hpsi failed: %d" % code)
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / steps
        ph = h.timing() if timed else None
        if timed:
            h.set_timing(False)
        return ms, ph

    # the step time with the launches overlapping as in production; then the
    # same steps with per-kernel events (launches serialised) for the table
    ms64, _ = run("f64", args.steps, False)
    tflops = 2.0 * h.macs * R / (ms64 * 1e-3) / 1e12
    ms64_serial, ph = run("f64", args.steps, True)
    # per-kernel figures from one event pair around every launch; the
    # dominant kernel is the one with the largest summed time, and its
    # algorithmic FLOPs per launch are the plan's per-kernel count (tiles of
    # that launch, all right-hand sides) divided by its launches per step
    def kernel_table(ph, kfl):
        kms, kn = ph
        table = {}
        for k in range(L.NKERNELS):
            if kn[k] == 0:
                continue
            per_step = kn[k] // args.steps
            avg = kms[k] / kn[k]
            fl = kfl[k] * R / max(per_step, 1)
            table[L.kernel_name(k)] = dict(
                launches_per_step=per_step, avg_launch_ms=round(avg, 4),
                tflops=round(fl / (avg * 1e-3) / 1e12, 3) if fl > 0 else None)
        dk = max((k for k in range(L.NKERNELS) if kn[k] and kfl[k] > 0),
                 key=lambda k: kms[k])
        d_avg = kms[dk] / kn[dk]
        d_ach = kfl[dk] * R / max(kn[dk] // args.steps, 1) / (d_avg * 1e-3) / 1e12
        return (table, L.kernel_name(dk), d_avg, kn[dk] // args.steps, d_ach,
                sum(kms[k] for k in range(L.NKERNELS)) / args.steps)

    kernels, dom, dur, per_step, achieved, t_gemm = kernel_table(
        ph, h.kernel_flops("f64"))
    log("FP64: %.3f ms/step, %.2f TFLOP/s; launches serialised for the "
        "per-kernel events: %.3f ms/step (%.3f ms in kernels)" % (
            ms64, tflops, ms64_serial, t_gemm))
    for name, kd in sorted(kernels.items()):
        log("  %-16s %2d launch/step %8.3f ms/launch %s TF" % (
            name, kd["launches_per_step"], kd["avg_launch_ms"],
            kd["tflops"]))
    ms32, _ = run("f32", max(3, args.steps // 2), False)
    tflops32 = 2.0 * h.macs * R / (ms32 * 1e-3) / 1e12
    log("FP32: %.3f ms/step, %.2f TFLOP/s" % (ms32, tflops32))
    msb, _ = run("bf16x3", max(3, args.steps // 2), False)
    tflopsb = 2.0 * h.macs * R / (msb * 1e-3) / 1e12
    _, phb = run("bf16x3", args.steps, True)
    kernels_b, dom_b, dur_b, _, ach_b, _ = kernel_table(
        phb, h.kernel_flops("bf16x3"))
    log("split-BF16 x3: %.3f ms/step, %.2f TFLOP/s (FP64-equivalent work); "
        "dominant %s %.3f ms/launch, %.1f TFLOP/s" % (
            msb, tflopsb, dom_b, dur_b, ach_b))

    # fixed-tile baseline (64x64 everywhere) on the same problem
    os.environ["B2G_TILE"] = "64"
    hf = EffectiveHamiltonian(pb, ctx, fp32=False, device_data=(ldev, rdev))
    os.environ.pop("B2G_TILE", None)
    hh, h = h, hf
    msf, _ = run("f64", max(3, args.steps // 2), False)
    h = hh
    hf.deallocate()
    tflops_fixed = 2.0 * h.macs * R / (msf * 1e-3) / 1e12
    log("FP64 fixed 64x64 tiles: %.3f ms/step, %.2f TFLOP/s" % (
        msf, tflops_fixed))

    extra = {}
    if not args.no_davidson:
        # DMRG-like initial guess: the lowest diagonal state plus noise
        d = h.diag
        guess = 0.3 * torch.randn(h.ket_len, generator=g, dtype=torch.float64,
                                  device=dev) / float(np.sqrt(h.ket_len))
        guess[int(torch.argmin(d))] += 1.0
        guess /= torch.linalg.norm(guess)
        # cold: a fresh effective Hamiltonian per solve, as at every site of
        # a sweep (graph capture, workspace and, for split-BF16, the packed
        # operator planes included); warm: a second solve on the same one
        t_create = {}

        def solve(prec, fresh):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            hh = (EffectiveHamiltonian(pb, ctx, fp32=True,
                                       device_data=(ldev, rdev))
                  if fresh else h)
            torch.cuda.synchronize()
            if fresh:
                t_create[prec] = time.perf_counter() - t0
            x = guess.clone()
            e, nd, _, t = hh.eigs(x, conv_thrd=1e-9, precision=prec,
                                  refine=prec != "f64")
            if fresh:
                hh.deallocate()
            return e, nd, t, x

        e64, nd64, t64c, _ = solve("f64", True)
        h.eigs(guess.clone(), conv_thrd=1e-9)
        e64, nd64, t64, x64 = solve("f64", False)
        dav = dict(conv_thrd=1e-9, e_fp64=e64, ndav_fp64=nd64,
                   s_fp64=round(t64, 4), s_fp64_cold=round(t64c, 4))
        log("Davidson FP64: E=%.12f ndav=%d %.4fs warm, %.4fs cold" % (
            e64, nd64, t64, t64c))
        # mixed precision: reduced-precision matvecs to convergence, then an
        # FP64 refinement from the converged vector
        for tag, prec in (("f32", "f32"), ("bf16x3", "bf16x3")):
            _, _, tmc, _ = solve(prec, True)
            h.eigs(guess.clone(), conv_thrd=1e-9, precision=prec,
                   refine=True)
            emx, ndmx, tmx, xmx = solve(prec, False)
            ov = float(torch.dot(x64, xmx).abs())
            dav["mixed_" + tag] = dict(
                e=emx, ndav=ndmx, s=round(tmx, 4), s_cold=round(tmc, 4),
                abs_dE=abs(emx - e64), infidelity=max(0.0, 1.0 - ov),
                speedup_vs_fp64=round(t64 / tmx, 3) if tmx > 0 else None,
                speedup_vs_fp64_cold=round(t64c / tmc, 3) if tmc > 0 else None)
            log("Davidson %s+FP64 refine: E=%.12f ndav=%d %.4fs warm "
                "(%.2fx), %.4fs cold (%.2fx) |dE|=%.2e 1-|<a|b>|=%.2e" % (
                    tag, emx, ndmx, tmx, t64 / tmx, tmc, t64c / tmc,
                    abs(emx - e64), max(0.0, 1.0 - ov)))
        extra["davidson"] = dav
        # FP64 truncation of the converged FP64 wavefunction on the device
        # (density matrix, block eigensolver, selection, split), forward
        # direction, k = M states, as a sweep does after the solve
        from b2g import truncate as TR
        o, n = int(pb.ket.off[0]), int(pb.ket.n[0])
        kinfo = TR.BlockInfo(pb.ket.quanta[o:o + n], pb.ket.nbra[o:o + n],
                             pb.ket.nket[o:o + n], pb.ket.shift[o:o + n],
                             int(pb.ket.dq[0]))

        def trunc():
            torch.cuda.synchronize()
            t = time.perf_counter()
            sp = TR.split_density_matrix(ctx, 1, kinfo, x64, args.m, 1e-14,
                                         True, True)
            torch.cuda.synchronize()
            return sp, time.perf_counter() - t
        trunc()
        sp, t_split = trunc()
        dmi, dmd = sp.dm
        mats = [dmd[int(dmi.shift[i]):int(dmi.shift[i]) +
                    int(dmi.nbra[i]) ** 2].view(int(dmi.nbra[i]), -1)
                for i in range(len(dmi))]
        torch.cuda.synchronize()
        t = time.perf_counter()
        _, _, sweeps = TR.eigh_batched(ctx, mats)
        torch.cuda.synchronize()
        t_eig = time.perf_counter() - t
        extra["truncation"] = dict(
            k=args.m, cutoff=1e-14, s=round(t_split, 4),
            s_eigensolver=round(t_eig, 4), jacobi_sweeps=sweeps,
            error=sp.error, dm_blocks=len(dmi),
            dm_states=int(dmi.nbra.sum()), max_dm_block=int(dmi.nbra.max()),
            kept=int(sp.left[0].nket.sum()))
        log("truncation (k=%d): %.4fs (eigensolver %.4fs, %d sweeps), "
            "%d density-matrix blocks, %d states, largest %d, error %.3e" % (
                args.m, t_split, t_eig, sweeps, len(dmi),
                int(dmi.nbra.sum()), int(dmi.nbra.max()), sp.error))
        trunc_host = (kinfo, x64.cpu().numpy())
        # one two-site step of a sweep at this site: a fresh effective
        # Hamiltonian (setup), the cold solve, the truncation; the middle
        # site is the largest of the chain, so 2 (n - 1) steps bound a
        # sweep from above
        steps = 2 * (args.n_orb - 1)
        site = {}
        for tag, tc, key in (("fp64", t64c, "f64"),
                             ("bf16x3", dav["mixed_bf16x3"]["s_cold"],
                              "bf16x3"),
                             ("f32", dav["mixed_f32"]["s_cold"], "f32")):
            st = t_create[key] + tc + t_split
            site[tag] = dict(setup_s=round(t_create[key], 4),
                             solve_s=round(tc, 4),
                             truncation_s=round(t_split, 4),
                             step_s=round(st, 4),
                             sweep_upper_bound_s=round(steps * st, 3))
        site["speedup_bf16x3_vs_fp64"] = round(
            site["fp64"]["step_s"] / site["bf16x3"]["step_s"], 3)
        extra["site_step"] = site
        log("site step (setup + cold solve + truncation): FP64 %.4fs, "
            "split-BF16+refine %.4fs (%.2fx); sweep bound (%d steps): "
            "%.2fs / %.2fs" % (site["fp64"]["step_s"],
                               site["bf16x3"]["step_s"],
                               site["speedup_bf16x3_vs_fp64"], steps,
                               site["fp64"]["sweep_upper_bound_s"],
                               site["bf16x3"]["sweep_upper_bound_s"]))

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        lh, rh = ldev.cpu().numpy(), rdev.cpu().numpy()

        def y_gpu(x):
            xd = torch.tensor(x, dtype=torch.float64, device=dev)
            return h.multiply(xd).cpu().numpy()

        try:
            cpu = cpu_reference(pb, lh, rh, y_gpu, threads)
        except Exception as e:  # reported, and the restatement stands in
            log("block2 CPU baseline failed:", repr(e))
            cpu = None
        log("CPU baseline (block2):", cpu)
        port = cpu_baseline(pb, lh, rh, args.cpu_seconds, threads)
        log("CPU restatement:", port)
        if cpu is None:
            cpu = port
        else:
            cpu["port"] = port
        del lh, rh
        if "truncation" in extra:
            # the same truncation by the oracle on the host (numpy / LAPACK)
            import truncate_oracle
            ki, xh = trunc_host
            oi = dict(quanta=ki.quanta, nbra=ki.nbra, nket=ki.nket,
                      shift=ki.shift, dq=ki.dq)
            t = time.perf_counter()
            r = truncate_oracle.split_density_matrix(1, oi, xh, args.m, 1e-14,
                                                     True, True, 0)
            t_cpu = time.perf_counter() - t
            extra["truncation"]["cpu_oracle_s"] = round(t_cpu, 4)
            extra["truncation"]["abs_error_vs_cpu"] = abs(r["error"] -
                                                          extra["truncation"]["error"])
            log("truncation on the host (numpy/LAPACK, %d threads): %.4fs" % (
                threads, t_cpu))

    out = {
        "metric": METRIC,
        "value": round(tflops, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms64, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "batched H*psi, middle-site two-site effective "
                        "Hamiltonian, n=%d orbitals, M=%d, SU(2), %d RHS, "
                        "seed %d" % (args.n_orb, args.m, R, args.seed),
            "psi_len": h.ket_len, "psi_blocks": int(len(pb.ket.quanta)),
            "max_block_dim": args.max_dim, "n_terms": int(len(pb.term_conj)),
            "n_pairs": h.n_pairs, "macs_per_rhs": h.macs,
            "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "kernel": dom,
                     "launches_per_step": per_step,
                     "avg_launch_ms": round(dur, 4),
                     "achieved": round(achieved, 4), "peak": FP64_MFMA_PEAK,
                     "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_MFMA_PEAK, 4),
                     "traffic": profiled_traffic(dom)},
        "kernels_fp64": kernels,
        "autotune": {"block_dim": rep, "tile_rates_tflops":
                     [round(r, 2) for r in rates],
                     "tile_rates_bf16x3": [round(r, 2) for r in rates_bf],
                     "measured": tuning,
                     "fixed_64x64_tflops": round(tflops_fixed, 4),
                     "speedup_vs_fixed": round(tflops / tflops_fixed, 3)},
        "fp32": {"value": round(tflops32, 4), "unit": "TFLOP/s",
                 "ms_per_step": round(ms32, 4),
                 "frac_of_fp32_peak": round(tflops32 / FP32_MFMA_PEAK, 4)},
        "bf16x3": {"value": round(tflopsb, 4), "unit": "TFLOP/s",
                   "ms_per_step": round(msb, 4),
                   "note": "FP64-equivalent FLOPs; 3 BF16 MFMA products "
                           "per multiply-add, FP32 accumulation",
                   "dominant_kernel": {
                       "kernel": dom_b, "avg_launch_ms": round(dur_b, 4),
                       "achieved_fp64_equiv_tflops": round(ach_b, 3),
                       "bf16_mfma_frac": round(3 * ach_b / 2500.0, 4),
                       "traffic": profiled_traffic("bf16x3_" + dom_b)},
                   "kernels": kernels_b},
        "cpu_baseline": cpu,
    }
    out.update(extra)
    if not args.no_sweep:
        out["sweep"] = sweep_leg(ctx, min(16, os.cpu_count() or 1),
                                 not args.no_cpu)
        out["sweep_abinitio"] = abinitio_leg([50, 100, 250, 500, 1000])
    print(json.dumps(out), flush=True)
    h.deallocate()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
