"""Density-matrix truncation on the device and its batched eigensolver.

split_density_matrix follows MovingEnvironment::split_density_matrix
(src/dmrg/moving_environment.hpp:4250-4345) after
MovingEnvironment::density_matrix (:3554-3577) for one wavefunction with no
noise, as DMRG::sweep calls them (src/dmrg/sweep_algorithm.hpp:430-433):
density matrix, blockwise eigendecomposition, selection of the kept states
(truncate_density_matrix, :3716-3853) and the two factors.  eigh_batched is
the device counterpart of GMatrixFunctions::eigs
(src/core/matrix_functions.hpp:1672-1688) over many blocks at once.
Everything runs in libb2g (FP64 on the GPU); this module only moves
arguments and results.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib as L

TRUNC_PHYSICAL = 0
TRUNC_REDUCED = 1
TRUNC_REDUCED_INVERSED = 2
TRUNC_REAL_DENSITY_MATRIX = 4
TRUNC_SPECTRA_WITH_MULTIPLICITY = 8
TRUNC_KEEP_ONE = 16
NOISE_WAVEFUNCTION = 1
NOISE_DENSITY_MATRIX = 2


class BlockInfo:
    """A block-sparse matrix structure (SparseMatrixInfo): quanta, row and
    column dimensions and element offsets of each block, delta quantum."""

    def __init__(self, quanta, nbra, nket, shift=None, dq=0, fermion=0):
        self.quanta = np.ascontiguousarray(quanta, dtype=np.uint64)
        self.nbra = np.ascontiguousarray(nbra, dtype=np.uint32)
        self.nket = np.ascontiguousarray(nket, dtype=np.uint32)
        if shift is None:
            sz = self.nbra.astype(np.uint64) * self.nket.astype(np.uint64)
            shift = np.concatenate([[0], np.cumsum(sz)[:-1]]) if len(sz) else []
        self.shift = np.ascontiguousarray(shift, dtype=np.uint64)
        self.dq = int(dq)
        self.fermion = int(fermion)

    def __len__(self):
        return len(self.quanta)

    @property
    def total(self):
        if len(self) == 0:
            return 0
        return int(self.shift[-1]) + int(self.nbra[-1]) * int(self.nket[-1])

    def block(self, data, i):
        o, m, n = int(self.shift[i]), int(self.nbra[i]), int(self.nket[i])
        return data[o:o + m * n].reshape(m, n)

    def to_c(self):
        return L.b2g_sparse_info(
            self.dq, self.fermion, len(self),
            L.ptr(self.quanta, C.c_uint64), L.ptr(self.nbra, C.c_uint32),
            L.ptr(self.nket, C.c_uint32), L.ptr(self.shift, C.c_uint64))

    @staticmethod
    def from_c(ci):
        n = ci.n
        if n == 0:
            return BlockInfo([], [], [], [], ci.delta_quantum, ci.is_fermion)
        return BlockInfo(np.ctypeslib.as_array(ci.quanta, (n,)).copy(),
                         np.ctypeslib.as_array(ci.n_states_bra, (n,)).copy(),
                         np.ctypeslib.as_array(ci.n_states_ket, (n,)).copy(),
                         np.ctypeslib.as_array(ci.block_shifts, (n,)).copy(),
                         ci.delta_quantum, ci.is_fermion)


def _dev_copy(ctx, ptr, n):
    """a torch tensor holding n doubles copied from device memory `ptr`"""
    out = torch.empty(n, dtype=torch.float64, device=ctx.device)
    if n:
        L.check(ctx.lib.b2g_vec_copy(ctx.h, C.c_void_p(ptr),
                                     C.c_void_p(out.data_ptr()), n))
    return out


def eigh_batched(ctx, mats):
    """Eigen-decomposition of real symmetric FP64 device matrices.

    Returns (ws, vs, sweeps): ws[i] ascending eigenvalues of mats[i], vs[i]
    its eigenvectors as rows, each with its largest-magnitude component
    positive."""
    n = np.array([m.shape[0] for m in mats], dtype=np.uint32)
    sizes = n.astype(np.uint64) ** 2
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.uint64) \
        if len(mats) else np.zeros(0, np.uint64)
    a = torch.cat([m.reshape(-1) for m in mats]) if len(mats) else \
        torch.zeros(1, dtype=torch.float64, device=ctx.device)
    a = a.contiguous()
    v = torch.empty_like(a)
    w = torch.empty(max(int(n.sum()), 1), dtype=torch.float64,
                    device=ctx.device)
    torch.cuda.synchronize(ctx.device)
    sw = C.c_int(0)
    L.check(ctx.lib.b2g_syev_batched(ctx.h, len(mats), L.ptr(n, C.c_uint32),
                                     L.ptr(off, C.c_uint64),
                                     C.c_void_p(a.data_ptr()),
                                     C.c_void_p(w.data_ptr()),
                                     C.c_void_p(v.data_ptr()), C.byref(sw)))
    ctx.synchronize()
    ws, vs, wo = [], [], 0
    for i in range(len(mats)):
        k = int(n[i])
        ws.append(w[wo:wo + k])
        vs.append(v[int(off[i]):int(off[i]) + k * k].view(k, k))
        wo += k
    return ws, vs, sw.value


class Split:
    """Result of split_density_matrix: error, left and right factors
    (BlockInfo + device tensor), the density matrix before
    diagonalisation, all eigenvalues (per density-matrix block, ascending)
    and the wavefunction spectra."""

    def __init__(self, error, left, right, dm, eigvals, spectra):
        self.error = error
        self.left, self.right, self.dm = left, right, dm
        self.eigvals = eigvals
        self.spectra = spectra


class Rng:
    """block2's random generator (std::mt19937 as Random::rand_seed(seed)
    leaves it) held by the library; noise drawn through it continues one
    stream across calls, as block2's generator does across a sweep."""

    def __init__(self, seed):
        self.lib = L.lib()
        h = C.c_void_p()
        L.check(self.lib.b2g_rng_create(int(seed), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.b2g_rng_destroy(self.h)
            self.h = None


def split_density_matrix(ctx, symmetry, info, psi, k, cutoff=0.0,
                         trace_right=True, normalize=True,
                         trunc_type=TRUNC_PHYSICAL, scale=1.0, weights=None,
                         noise=0.0, noise_type=NOISE_WAVEFUNCTION, seed=0,
                         rng=None, extra=None):
    """Truncate the wavefunction psi (device FP64 tensor with the layout of
    `info`) to k states (-1: all) through its density matrix.  A 2-d psi
    holds state-averaged roots (one per row, density matrix weighted by
    `weights`, multi_split_density_matrix); the wavefunction factor then
    comes back with one row per root.  noise > 0 adds block2's
    wavefunction or density-matrix noise (for each root), drawn from
    std::mt19937 seeded with `seed` as block2's Random after
    Random::rand_seed(seed).  Each call starts a new generator: with noise
    the seed must be nonzero (block2 takes 0 as a clock seed) and a sweep
    passes a new one at each site; or pass `rng` (an Rng) to draw from one
    continuing stream.  extra: [(BlockInfo, device vector)] added to the
    density matrix as v v^T (v^T v backward), the perturbed wavefunctions
    of perturbative noise."""
    lib = ctx.lib
    ci = info.to_c()
    psi = psi.contiguous()
    torch.cuda.synchronize(ctx.device)
    h = C.c_void_p()
    if extra:
        if noise or psi.dim() != 1:
            raise ValueError("extra terms with one root and no other noise")
        parts, offs, t = [psi], [], psi.numel()
        for _, v in extra:
            offs.append(t)
            parts.append(v.reshape(-1))
            t += v.numel()
        buf = torch.cat(parts)
        xinf = (L.b2g_sparse_info * len(extra))(*[x.to_c() for x, _ in extra])
        xoff = np.ascontiguousarray(offs, dtype=np.uint64)
        L.check(lib.b2g_split_density_matrix_extra(
            ctx.h, int(symmetry), C.byref(ci), C.c_void_p(buf.data_ptr()),
            len(extra), xinf, L.ptr(xoff, C.c_uint64), int(k), float(cutoff),
            int(bool(trace_right)), int(bool(normalize)), int(trunc_type),
            float(scale), C.byref(h)))
    elif noise and rng is not None:
        nr = 1 if psi.dim() == 1 else psi.shape[0]
        w = np.ascontiguousarray(
            np.full(nr, 1.0 / nr) if weights is None else weights,
            dtype=np.float64)
        assert w.size == nr
        L.check(lib.b2g_split_density_matrix_noise_rng(
            ctx.h, int(symmetry), C.byref(ci), C.c_void_p(psi.data_ptr()),
            nr, L.ptr(w, C.c_double), float(noise), int(noise_type),
            rng.h, int(k), float(cutoff), int(bool(trace_right)),
            int(bool(normalize)), int(trunc_type), float(scale), C.byref(h)))
    elif noise:
        nr = 1 if psi.dim() == 1 else psi.shape[0]
        w = np.ascontiguousarray(
            np.full(nr, 1.0 / nr) if weights is None else weights,
            dtype=np.float64)
        assert w.size == nr
        L.check(lib.b2g_split_density_matrix_noise(
            ctx.h, int(symmetry), C.byref(ci), C.c_void_p(psi.data_ptr()),
            nr, L.ptr(w, C.c_double), float(noise), int(noise_type),
            int(seed), int(k), float(cutoff), int(bool(trace_right)),
            int(bool(normalize)), int(trunc_type), float(scale), C.byref(h)))
    elif psi.dim() == 1:
        L.check(lib.b2g_split_density_matrix(
            ctx.h, int(symmetry), C.byref(ci), C.c_void_p(psi.data_ptr()),
            int(k), float(cutoff), int(bool(trace_right)),
            int(bool(normalize)), int(trunc_type), float(scale), C.byref(h)))
    else:
        nr = psi.shape[0]
        w = np.ascontiguousarray(
            np.full(nr, 1.0 / nr) if weights is None else weights,
            dtype=np.float64)
        assert w.size == nr
        L.check(lib.b2g_split_density_matrix_multi(
            ctx.h, int(symmetry), C.byref(ci), C.c_void_p(psi.data_ptr()),
            nr, L.ptr(w, C.c_double), int(k), float(cutoff),
            int(bool(trace_right)), int(bool(normalize)), int(trunc_type),
            float(scale), C.byref(h)))
    try:
        nroots = C.c_int(1)
        L.check(lib.b2g_split_n_roots(h, C.byref(nroots)))
        wav_part = 1 if trace_right else 0
        parts = []
        for which in range(3):
            ri = L.b2g_sparse_info()
            dp = C.c_void_p()
            n = C.c_size_t()
            L.check(lib.b2g_split_result(h, which, C.byref(ri), C.byref(dp),
                                         C.byref(n)))
            if which == wav_part and psi.dim() == 2:
                data = _dev_copy(ctx, dp.value,
                                 n.value * nroots.value).view(nroots.value, -1)
            else:
                data = _dev_copy(ctx, dp.value, n.value)
            parts.append((BlockInfo.from_c(ri), data))
        err = C.c_double()
        L.check(lib.b2g_split_error(h, C.byref(err)))
        vals = []
        for which in range(2):
            n = C.c_size_t(0)
            L.check(lib.b2g_split_values(h, which, None, C.byref(n)))
            a = np.zeros(max(n.value, 1))
            L.check(lib.b2g_split_values(h, which, L.ptr(a, C.c_double),
                                         C.byref(n)))
            vals.append(a[:n.value])
        ctx.synchronize()
    finally:
        lib.b2g_split_destroy(h)
    return Split(err.value, parts[0], parts[1], parts[2], vals[0], vals[1])
