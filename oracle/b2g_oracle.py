"""ORACLE -- TEST INFRASTRUCTURE ONLY.

ctypes binding of oracle/_build/libb2g_oracle.so, the plain-C CPU restatement
of block2's effective-Hamiltonian hot path (see b2g_oracle.h for the
reference file:line each function follows).  Imported only by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libb2g_oracle.so")

_u64 = C.POINTER(C.c_uint64)
_u32 = C.POINTER(C.c_uint32)
_i32 = C.POINTER(C.c_int32)
_u8 = C.POINTER(C.c_uint8)
_f64 = C.POINTER(C.c_double)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: make -C oracle")
        L = C.CDLL(LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = (
            [C.c_int, C.c_uint64, C.c_uint64]
            + [C.c_uint64, C.c_int, _u64, _u32, _u32, _u64] * 2
            + [C.c_int, _u64, _u32, _u32, _u32, _u64, _u32, _u32, _u64] * 2
            + [C.c_int, _f64, _i32, _u64, _f64] * 2
            + [C.c_int, _u8, _i32, _i32, _f64])
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_error.restype = C.c_char_p
        L.orc_error.argtypes = [C.c_void_p]
        for f in ("orc_ket_len", "orc_bra_len", "orc_macs"):
            getattr(L, f).restype = C.c_uint64
            getattr(L, f).argtypes = [C.c_void_p]
        L.orc_connections.restype = C.c_int64
        L.orc_connections.argtypes = [C.c_void_p, C.c_int, _u32, _u32, _u32,
                                      _u64, _f64, _i32, C.POINTER(C.c_int64),
                                      _u64, _u32]
        L.orc_hpsi.argtypes = [C.c_void_p, _f64, _f64, C.c_int]
        L.orc_hpsi_sample.restype = C.c_uint64
        L.orc_hpsi_sample.argtypes = [C.c_void_p, _f64, _f64, C.c_uint64,
                                      C.c_int]
        L.orc_diag.argtypes = [C.c_void_p, _f64]
        L.orc_hpsi_blocks.restype = C.c_uint64
        L.orc_hpsi_blocks.argtypes = [C.c_void_p, _f64, _f64, C.c_int,
                                      _i32, C.c_int]
        L.orc_block_macs.argtypes = [C.c_void_p, _u64]
        L.orc_davidson.restype = C.c_int
        L.orc_davidson.argtypes = [C.c_void_p, _f64, C.c_int, C.c_double,
                                   C.c_double, C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.c_int, C.c_double, C.c_int,
                                   C.c_void_p, C.c_void_p, _f64,
                                   C.POINTER(C.c_int), C.c_int]
        L.orc_grad_ops.argtypes = [C.c_void_p, _f64, _f64, _f64, _f64]
        L.orc_sixj.restype = C.c_double
        L.orc_sixj.argtypes = [C.c_int] * 6
        L.orc_ninej.restype = C.c_double
        L.orc_ninej.argtypes = [C.c_int] * 9
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


class Oracle:
    """CPU restatement bound to one Problem (b2g.problem.Problem)."""

    def __init__(self, pb):
        L = lib()
        self.pb = pb
        k, b, li, ri, lo, ro = pb.ket, pb.bra, pb.linfo, pb.rinfo, pb.lops, pb.rops

        def info_args(s):
            return [int(s.dq[0]), int(s.n[0]), _p(s.quanta, _u64),
                    _p(s.nbra, _u32), _p(s.nket, _u32), _p(s.shift, _u64)]

        def infos_args(s):
            return [len(s), _p(s.dq, _u64), _p(s.fermion, _u32), _p(s.n, _u32),
                    _p(s.off, _u32), _p(s.quanta, _u64), _p(s.nbra, _u32),
                    _p(s.nket, _u32), _p(s.shift, _u64)]

        def ops_args(o):
            return [len(o), _p(o.factor, _f64), _p(o.info, _i32),
                    _p(o.off, _u64), _p(o.data, _f64)]

        args = ([pb.sym, pb.opdq, pb.hop_left_vacuum] + info_args(k)
                + info_args(b) + infos_args(li) + infos_args(ri) + ops_args(lo)
                + ops_args(ro) + [len(pb.term_conj), _p(pb.term_conj, _u8),
                                  _p(pb.term_a, _i32), _p(pb.term_b, _i32),
                                  _p(pb.term_factor, _f64)])
        self.h = L.orc_create(*args)
        self.ket_len = L.orc_ket_len(self.h)
        self.bra_len = L.orc_bra_len(self.h)
        self.macs = L.orc_macs(self.h)

    def __del__(self):
        try:
            lib().orc_destroy(self.h)
        except Exception:
            pass

    def connections(self, which=0):
        L = lib()
        n = L.orc_connections(self.h, which, None, None, None, None, None,
                              None, None, None, None)
        r = dict(ia=np.zeros(n, np.uint32), ib=np.zeros(n, np.uint32),
                 ic=np.zeros(n, np.uint32), stride=np.zeros(n, np.uint64),
                 factor=np.zeros(n, np.float64), n=np.zeros(5, np.int32))
        nq = C.c_int64()
        L.orc_connections(self.h, which, None, None, None, None, None, None,
                          C.byref(nq), None, None)
        r["quanta"] = np.zeros(nq.value, np.uint64)
        r["idx"] = np.zeros(nq.value, np.uint32)
        L.orc_connections(self.h, which, _p(r["ia"], _u32), _p(r["ib"], _u32),
                          _p(r["ic"], _u32), _p(r["stride"], _u64),
                          _p(r["factor"], _f64), _p(r["n"], _i32), C.byref(nq),
                          _p(r["quanta"], _u64), _p(r["idx"], _u32))
        return r

    def hpsi(self, x, nthreads=1):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.bra_len)
        lib().orc_hpsi(self.h, _p(x, _f64), _p(y, _f64), nthreads)
        return y

    def hpsi_sample(self, x, max_pairs, nthreads=1):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.bra_len)
        macs = lib().orc_hpsi_sample(self.h, _p(x, _f64), _p(y, _f64),
                                     int(max_pairs), nthreads)
        return y, macs

    def hpsi_blocks(self, x, blocks, nthreads=1):
        """H x on the listed output blocks only (others zero); returns
        (y, multiply-adds)."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.bra_len)
        b = np.ascontiguousarray(blocks, np.int32)
        macs = lib().orc_hpsi_blocks(self.h, _p(x, _f64), _p(y, _f64),
                                     len(b), _p(b, _i32), int(nthreads))
        return y, macs

    def block_macs(self):
        m = np.zeros(int(self.pb.bra.n[0]), np.uint64)
        lib().orc_block_macs(self.h, _p(m, _u64))
        return m

    def diag(self):
        d = np.zeros(self.ket_len)
        lib().orc_diag(self.h, _p(d, _f64))
        return d

    def davidson(self, x, conv=1e-10, max_iter=500, def_min=2, def_max=50,
                 rel_conv=0.0, soft_max_iter=-1, davidson_type=0, shift=0.0,
                 ors=None, proj_weights=None, nthreads=1):
        """Returns (energies, Ritz vectors, ndav, rc); rc 0 converged, 1
        max_iter, 2 soft_max_iter."""
        x = np.array(x, np.float64, copy=True, order="C")
        k = 1 if x.ndim == 1 else x.shape[0]
        e = np.zeros(k)
        nd = C.c_int()
        nor, op, wp = 0, None, None
        if ors is not None and len(ors):
            ors = np.ascontiguousarray(np.atleast_2d(ors), np.float64)
            nor, op = ors.shape[0], ors.ctypes.data
            if proj_weights is not None:
                proj_weights = np.ascontiguousarray(proj_weights, np.float64)
                assert len(proj_weights) == nor
                wp = proj_weights.ctypes.data
        rc = lib().orc_davidson(self.h, _p(x, _f64), k, conv, rel_conv,
                                max_iter, soft_max_iter, def_min, def_max,
                                davidson_type, shift, nor, op, wp,
                                _p(e, _f64), C.byref(nd), int(nthreads))
        return e, x, nd.value, rc

    def grad_ops(self, v, c):
        gl = np.zeros(len(self.pb.lops.data))
        gr = np.zeros(len(self.pb.rops.data))
        v = np.ascontiguousarray(v, np.float64)
        c = np.ascontiguousarray(c, np.float64)
        lib().orc_grad_ops(self.h, _p(v, _f64), _p(c, _f64), _p(gl, _f64),
                           _p(gr, _f64))
        return gl, gr


def sixj(*a):
    return lib().orc_sixj(*a)


def ninej(*a):
    return lib().orc_ninej(*a)
