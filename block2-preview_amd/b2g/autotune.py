"""Tile autotuner for the grouped GEMM.

Measures the throughput of the nine workgroup tile shapes (TM, TN in
{32, 64, 128}) of the FP64 kernels, or of the split-BF16 kernels, and
installs the rates with b2g_set_tile_rates_for, so that the tiling of every
plan uploaded afterwards (FP64 / FP32 kernels), or of every packed
split-BF16 plan built afterwards, minimises its estimated time.  Two
measurements:

* `tune`: uniform block problems whose block size is the
  multiply-add-weighted median output-block size of the problem to be
  solved;
* `tune_on_problem`: the problem itself, tiled with one shape at a time,
  each shape's launches timed with per-kernel HIP events: the rate a shape
  reaches on this sector-size histogram (shallow segments and ragged edges
  included).

`choose` times the problem's product under each candidate rate table and
installs the faster.  Measurement goes through the C ABI with HIP events
on the library's stream.
"""
import ctypes as C
import os

import numpy as np
import torch

from . import _lib as L

SIZES = (32, 64, 128)


def representative_dim(pairs):
    """MAC-weighted median of min(v_m, v_n) over a pair list."""
    vm = pairs["v_m"].astype(np.float64)
    vn = pairs["v_n"].astype(np.float64)
    cm = pairs["c_m"].astype(np.float64)
    w = vm * cm * vn
    d = np.minimum(vm, vn)
    o = np.argsort(d)
    cw = np.cumsum(w[o])
    if cw[-1] <= 0:
        return 128
    return int(d[o][np.searchsorted(cw, 0.5 * cw[-1])])


def _measure(ctx, d, cfg, prec=0, target_macs=5e10, reps=3):
    tm, tn = 32 << (cfg // 3), 32 << (cfg % 3)
    lib = L.lib()
    blk = d * d
    n = max(4, int(target_macs / (2.0 * d ** 3)))
    lb, rb = C.c_void_p(), C.c_void_p()
    L.check(lib.b2g_buf_create(ctx, blk * 4, 0, C.byref(lb)))
    L.check(lib.b2g_buf_create(ctx, blk * 4, 0, C.byref(rb)))
    pairs = (L.b2g_pair * n)()
    yb = (L.b2g_block * n)()
    for i in range(n):
        pairs[i] = L.b2g_pair(a_off=(i % 4) * blk, b_off=((i * 3) % 4) * blk,
                              c_off=i * blk, v_off=i * blk, c_m=d, c_n=d,
                              v_m=d, v_n=d, alpha=1.0)
        yb[i] = L.b2g_block(i * blk, d, d)
    old = os.environ.get("B2G_TILE")
    os.environ["B2G_TILE"] = "%dx%d" % (tm, tn)
    plan = C.c_void_p()
    try:
        L.check(lib.b2g_upload_plan(ctx, lb, rb, pairs, n, yb, n, n * blk,
                                    n * blk, 1 if prec == 1 else 0,
                                    C.byref(plan)))
        # the packed split-BF16 plan is tiled when it is built: build it
        # while the tile shape is forced
        L.check(lib.b2g_plan_reserve(plan, 1, prec))
    finally:
        if old is None:
            os.environ.pop("B2G_TILE", None)
        else:
            os.environ["B2G_TILE"] = old
    x = torch.rand(n * blk, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    L.check(lib.b2g_hpsi(ctx, plan, x.data_ptr(), y.data_ptr(), 1, 1.0, 0,
                         prec))
    L.check(lib.b2g_ctx_synchronize(ctx))
    stream = torch.cuda.ExternalStream(lib.b2g_ctx_stream(ctx))
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        L.check(lib.b2g_hpsi(ctx, plan, x.data_ptr(), y.data_ptr(), 1, 1.0,
                             0, prec))
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    lib.b2g_plan_destroy(plan)
    lib.b2g_buf_destroy(lb)
    lib.b2g_buf_destroy(rb)
    return 4.0 * n * d ** 3 / (ms * 1e-3) / 1e12


def tune(ctx, dim=256, install=True, precision="f64"):
    """Measure the nine tile shapes at block size `dim` with the kernels of
    `precision` ("f64" or "bf16x3"); returns TFLOP/s (FP64-equivalent) per
    configuration and (optionally) installs them as that precision's tiling
    rates."""
    h = ctx.h if hasattr(ctx, "h") else ctx
    prec = L.PREC[precision]
    rates = [_measure(h, int(dim), cfg, prec) for cfg in range(9)]
    if install:
        arr = (C.c_double * 9)(*rates)
        L.check(L.lib().b2g_set_tile_rates_for(prec, arr, 9))
    return rates


def _forced(shape):
    """Context manager: B2G_TILE forced to `shape` (None: unset)."""
    class _F:
        def __enter__(self):
            self.old = os.environ.get("B2G_TILE")
            if shape is None:
                os.environ.pop("B2G_TILE", None)
            else:
                os.environ["B2G_TILE"] = shape
            return self

        def __exit__(self, *a):
            if self.old is None:
                os.environ.pop("B2G_TILE", None)
            else:
                os.environ["B2G_TILE"] = self.old
    return _F()


def _product_ms(h, x, precision, reps):
    """Average ms of `reps` products after one warm-up (which builds the
    packed split-BF16 plan)."""
    lib = L.lib()
    stream = torch.cuda.ExternalStream(lib.b2g_ctx_stream(h.ctx.h))
    h.multiply(x, precision=precision)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    y = torch.empty((x.shape[0], h.bra_len), dtype=torch.float64,
                    device=x.device)
    e0.record(stream)
    for _ in range(reps):
        L.check(lib.b2g_effham_hpsi(h.h, x.data_ptr(), y.data_ptr(),
                                    x.shape[0], 1.0, 0, L.PREC[precision]))
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_on_problem(ctx, problem, device_data, precision="f64", x=None,
                    reps=2, install=True):
    """Per-shape throughput (TFLOP/s, FP64-equivalent) of the nine tile
    shapes on `problem` itself: the product tiled with one shape at a time,
    the shape's phase-1 and phase-2 launches timed with per-kernel events;
    rate = their algorithmic FLOPs / their time.  x: (n_rhs, |psi|) device
    vectors (default one seeded vector)."""
    from .effham import EffectiveHamiltonian
    rates = []
    for cfg in range(9):
        shape = "%dx%d" % (32 << (cfg // 3), 32 << (cfg % 3))
        with _forced(shape):
            h = EffectiveHamiltonian(problem, ctx,
                                     fp32=(precision == "f32"),
                                     device_data=device_data)
            try:
                if x is None:
                    g = torch.Generator(device=ctx.device).manual_seed(7)
                    x = torch.randn((1, h.ket_len), generator=g,
                                    dtype=torch.float64, device=ctx.device)
                h.multiply(x, precision=precision)   # packs while forced
                h.set_timing(True)
                for _ in range(reps):
                    h.multiply(x, precision=precision)
                ms, n = h.timing()
                h.set_timing(False)
                fl = h.kernel_flops(precision)
            finally:
                h.deallocate()
        k1, k2 = L.K_PHASE1 + cfg, L.K_PHASE2 + cfg
        t = (ms[k1] + ms[k2]) * 1e-3
        f = (fl[k1] + fl[k2]) * x.shape[0] * reps
        rates.append(f / t / 1e12 if t > 0 and f > 0 else 0.0)
    # a shape without work on this histogram gets the slowest measured rate
    lo = min([r for r in rates if r > 0] or [1.0])
    rates = [r if r > 0 else lo for r in rates]
    if install:
        L.check(L.lib().b2g_set_tile_rates_for(
            L.PREC[precision], (C.c_double * 9)(*rates), 9))
    return rates


def choose(ctx, problem, device_data, tables, precision="f64", x=None,
           reps=3):
    """Times the product of `problem` tiled with each rate table of
    `tables` and installs the fastest; returns (index, ms per product of
    every table)."""
    from .effham import EffectiveHamiltonian
    lib = L.lib()
    times = []
    for rates in tables:
        L.check(lib.b2g_set_tile_rates_for(L.PREC[precision],
                                           (C.c_double * 9)(*rates), 9))
        with _forced(None):
            h = EffectiveHamiltonian(problem, ctx,
                                     fp32=(precision == "f32"),
                                     device_data=device_data)
            try:
                if x is None:
                    g = torch.Generator(device=ctx.device).manual_seed(7)
                    x = torch.randn((1, h.ket_len), generator=g,
                                    dtype=torch.float64, device=ctx.device)
                times.append(_product_ms(h, x, precision, reps))
            finally:
                h.deallocate()
    best = min(range(len(tables)), key=lambda i: times[i])
    L.check(lib.b2g_set_tile_rates_for(L.PREC[precision],
                                       (C.c_double * 9)(*tables[best]), 9))
    return best, times

