"""Tile autotuner for the grouped GEMM.

Measures the throughput of the nine workgroup tile shapes (TM, TN in
{32, 64, 128}) of the FP64 kernels, or of the split-BF16 kernels, on uniform
block problems whose block size is the multiply-add-weighted median
output-block size of the problem to be solved (its sector-size histogram),
and installs the rates with b2g_set_tile_rates_for, so that the tiling of
every plan uploaded afterwards (FP64 / FP32 kernels), or of every packed
split-BF16 plan built afterwards, minimises its estimated time.  Measurement goes through the thin C ABI with
HIP events on the library's stream.
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
