"""EffectiveHamiltonian on the GPU, mirroring block2's interface.

block2 (src/dmrg/effective_hamiltonian.hpp:99-1162, bound to Python at
src/pybind/pybind_dmrg.hpp:588-647) exposes precompute(), __call__ /
operator()(b, c, idx, factor, all_reduce), eigs(...) -> (energy, ndav, nflop,
tdav) and deallocate().  This class keeps those names and argument meanings;
vectors are torch float64 tensors on the context's device (PyTorch is only
the device-memory plumbing: every operation runs in libb2g.so).
"""
import ctypes as C

import numpy as np

import torch

from . import _lib as L


class Context:
    """A device context (one HIP stream) of libb2g.  Objects created on it
    (effective Hamiltonians) are released before the context itself."""

    def __init__(self, device=0):
        import weakref
        self.lib = L.lib()
        h = C.c_void_p()
        L.check(self.lib.b2g_ctx_create(int(device), 0, C.byref(h)))
        self.h = h
        self.device = torch.device("cuda", int(device))
        self._children = weakref.WeakSet()

    def synchronize(self):
        L.check(self.lib.b2g_ctx_synchronize(self.h))

    @property
    def num_cu(self):
        n = C.c_int()
        L.check(self.lib.b2g_ctx_num_cu(self.h, C.byref(n)))
        return n.value

    def close(self):
        if self.h:
            for child in list(self._children):
                child.deallocate()
            self.lib.b2g_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _dptr(t):
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise ValueError("expected a contiguous float64 CUDA tensor")
    return C.c_void_p(t.data_ptr())


class EffectiveKernel:
    """Hook around every Davidson matvec, as block2's EffectiveKernel
    (src/dmrg/effective_hamiltonian.hpp:83-93; Python trampoline
    src/pybind/pybind_dmrg.hpp:1107-1160).  compute(beta, f, a, b, xs)
    replaces b = beta * H a; f(a, b, scale) is the library's matvec on
    float64 device tensors.  The default calls f(a, b, beta).  Install with
    EffectiveHamiltonian.eff_kernel = kernel."""

    def compute(self, beta, f, a, b, xs):
        f(a, b, beta)


class EffectiveHamiltonian:
    """Two-site effective Hamiltonian of one DMRG site on the GPU.

    The constructor does what block2's constructor and precompute() do:
    connection lists (ConnectionInfo::initialize_wfn / initialize_diag),
    the diagonal, and the recorded pair list, uploaded once.
    """

    def __init__(self, problem, ctx, fp32=True, device_data=None):
        """device_data: optional (left, right) float64 CUDA tensors holding
        the operator buffers (e.g. from synth.fill_device)."""
        self.lib = L.lib()
        self.ctx = ctx
        self.problem = problem
        h = C.c_void_p()
        flags = L.FLAG_F32 if fp32 else 0
        if device_data is not None:
            ld, rd = device_data
            flags |= L.FLAG_DEVICE_DATA
            cp = problem.to_c((ld.data_ptr(), ld.numel(), rd.data_ptr(),
                               rd.numel()))
        else:
            cp = problem.to_c()
        L.check(self.lib.b2g_effham_create(ctx.h, C.byref(cp), flags,
                                           C.byref(h)))
        self.h = h
        ctx._children.add(self)
        kl, bl, npairs, macs = (C.c_uint64(), C.c_uint64(), C.c_uint64(),
                                C.c_uint64())
        L.check(self.lib.b2g_effham_info(h, C.byref(kl), C.byref(bl),
                                         C.byref(npairs), C.byref(macs)))
        self.ket_len, self.bra_len = kl.value, bl.value
        self.n_pairs, self.macs = npairs.value, macs.value
        self._diag = None

    # block2: precompute() records the GEMM list; here it is built and
    # uploaded by the constructor already.
    def precompute(self):
        return None

    def post_precompute(self):
        return None

    def __call__(self, b, c, idx=0, factor=1.0, all_reduce=True,
                 precision="f64"):
        """c += factor * H b (operator()(b, c, idx, factor, all_reduce),
        effective_hamiltonian.hpp:459-477).  b, c: (n,) or (n_rhs, n)."""
        if idx != 0:
            raise ValueError("only the ground operator (idx 0) is held")
        nrhs = self._check(b, c)
        torch.cuda.current_stream(self.ctx.device).synchronize()
        L.check(self.lib.b2g_effham_hpsi(self.h, _dptr(b), _dptr(c), nrhs,
                                         float(factor), 1, L.PREC[precision]))
        self.ctx.synchronize()

    def multiply(self, b, precision="f64", scale=1.0):
        """Return H b (overwrite semantics)."""
        nrhs = 1 if b.dim() == 1 else b.shape[0]
        c = torch.empty((nrhs, self.bra_len) if b.dim() > 1 else
                        (self.bra_len,), dtype=torch.float64, device=b.device)
        self._check(b, c)
        torch.cuda.current_stream(self.ctx.device).synchronize()
        L.check(self.lib.b2g_effham_hpsi(self.h, _dptr(b), _dptr(c), nrhs,
                                         float(scale), 0, L.PREC[precision]))
        self.ctx.synchronize()
        return c

    def _check(self, b, c):
        """Number of right-hand sides; rejects tensors whose sizes do not
        match the plan (the C ABI cannot see device allocation sizes)."""
        nrhs = 1 if b.dim() == 1 else b.shape[0]
        if b.dim() > 2 or b.numel() != nrhs * self.ket_len:
            raise ValueError("b must hold n_rhs x %d values" % self.ket_len)
        if c.numel() != nrhs * self.bra_len:
            raise ValueError("c must hold n_rhs x %d values" % self.bra_len)
        if b.device != self.ctx.device or c.device != self.ctx.device:
            raise ValueError("tensors must live on the context's device")
        return nrhs

    @property
    def plan(self):
        return self.lib.b2g_effham_plan(self.h)

    def set_timing(self, enable=True):
        L.check(self.lib.b2g_plan_set_timing(self.plan, 1 if enable else 0))

    def timing(self):
        """Per kernel id (b2g.h, L.kernel_name): (summed ms, launches)
        recorded since the last call, as two lists of L.NKERNELS."""
        ms = (C.c_double * L.NKERNELS)()
        n = (C.c_uint64 * L.NKERNELS)()
        L.check(self.lib.b2g_plan_timing(self.plan, ms, n))
        return list(ms), list(n)

    def kernel_flops(self, precision="f64"):
        """Algorithmic FLOPs of one H*psi per right-hand side, per kernel
        id (list of L.NKERNELS), for the launches of `precision` (the
        split-BF16 plan is tiled with its own tile rates)."""
        f = (C.c_double * L.NKERNELS)()
        L.check(self.lib.b2g_plan_kernel_flops_for(self.plan,
                                                   L.PREC[precision], f))
        return list(f)

    def pairs(self):
        """Recorded pair list as a numpy structured array."""
        import numpy as np
        n = C.c_uint64()
        L.check(self.lib.b2g_effham_pairs(self.h, C.byref(n), None))
        arr = (L.b2g_pair * max(n.value, 1))()
        L.check(self.lib.b2g_effham_pairs(self.h, C.byref(n), arr))
        return np.ctypeslib.as_array(arr)[:n.value]

    @property
    def diag(self):
        if self._diag is None:
            d = torch.empty(self.ket_len, dtype=torch.float64,
                            device=self.ctx.device)
            torch.cuda.current_stream(self.ctx.device).synchronize()
            L.check(self.lib.b2g_effham_diag(self.h, _dptr(d)))
            self.ctx.synchronize()
            self._diag = d
        return self._diag

    def eigs(self, ket, iprint=False, conv_thrd=5E-6, rel_conv_thrd=0.0,
             max_iter=5000, soft_max_iter=-1, deflation_min_size=2,
             deflation_max_size=50, davidson_type="Normal", shift=0.0,
             ortho_bra=None, projection_weights=None, precision="f64",
             refine=True):
        """Davidson for the targeted roots (eigs,
        effective_hamiltonian.hpp:480-568; k roots as the MultiMPS eigs,
        :1708-1800).  ket: (n,) or (k, n) CUDA tensor, overwritten with the
        Ritz vector(s).  davidson_type: a DavidsonTypes name or bit mask
        ("Normal", "GreaterThan", "LessThan", "CloseTo", "DavidsonPrecond",
        "NoPrecond", "Exact", "ElementProj", or e.g. "CloseTo|NoPrecond").
        ortho_bra: CUDA tensors of ket length (orthogonal states), with
        projection_weights optional.  Returns (energy or energies, ndav,
        nflop, tdav) with nflop in the reference's multiply-add
        convention."""
        k = 1 if ket.dim() == 1 else ket.shape[0]
        if ket.numel() != k * self.ket_len:
            raise ValueError("ket must hold k x %d values" % self.ket_len)
        if isinstance(davidson_type, str):
            t = 0
            for part in davidson_type.split("|"):
                if part.strip() not in L.DAV:
                    raise ValueError("unsupported Davidson type %r" % part)
                t |= L.DAV[part.strip()]
            davidson_type = t
        o = L.b2g_davidson_opts()
        self.lib.b2g_davidson_opts_default(C.byref(o))
        o.conv_thrd, o.rel_conv_thrd = float(conv_thrd), float(rel_conv_thrd)
        o.shift = float(shift)
        o.max_iter, o.soft_max_iter = int(max_iter), int(soft_max_iter)
        o.def_min, o.def_max = int(deflation_min_size), int(deflation_max_size)
        o.davidson_type = int(davidson_type)
        o.precision, o.refine = L.PREC[precision], 1 if refine else 0
        keep = []
        if ortho_bra:
            ors = [v.contiguous() for v in ortho_bra]
            for v in ors:
                if v.numel() != self.ket_len or v.dtype != torch.float64:
                    raise ValueError("ortho_bra vectors: %d FP64 values" %
                                     self.ket_len)
            arr = (C.c_void_p * len(ors))(*[v.data_ptr() for v in ors])
            keep += [ors, arr]
            o.n_ortho, o.ortho = len(ors), C.cast(arr, C.POINTER(C.c_void_p))
            if projection_weights is not None:
                w = np.ascontiguousarray(projection_weights, np.float64)
                if w.size != len(ors):
                    raise ValueError("one projection weight per ortho_bra")
                keep.append(w)
                o.proj_weights = L.ptr(w, C.c_double)
        en = (C.c_double * k)()
        nd, nfl, td = C.c_int(), C.c_uint64(), C.c_double()
        torch.cuda.current_stream(self.ctx.device).synchronize()
        code = self.lib.b2g_effham_eigs_ex(
            self.h, _dptr(ket), k, C.byref(o), en, C.byref(nd),
            C.byref(nfl), C.byref(td))
        self.ctx.synchronize()
        L.check(code)
        e = [en[i] for i in range(k)]
        return (e[0] if ket.dim() == 1 else e), nd.value, nfl.value, td.value

    def connections(self, which=0):
        """Host copy of the connection list (which 0: wfn, 1: diag)."""
        return connections(self.h, which)

    @property
    def eff_kernel(self):
        return getattr(self, "_kernel", None)

    @eff_kernel.setter
    def eff_kernel(self, kernel):
        """Install (or with None remove) an EffectiveKernel hook used by
        eigs (the host-driven Davidson loop runs while one is set)."""
        self._kernel = kernel
        if kernel is None:
            self._kernel_cb = None
            L.check(self.lib.b2g_effham_set_kernel(self.h, None, None))
            return
        ctx, lib = self.ctx, self.lib

        def trampoline(user, beta, f, f_data, a_ptr, b_ptr, n):
            try:
                dv = ctx.device
                a = torch.empty(n, dtype=torch.float64, device=dv)
                b = torch.zeros(n, dtype=torch.float64, device=dv)
                L.check(lib.b2g_vec_copy(ctx.h, C.c_void_p(a_ptr), _dptr(a), n))
                ctx.synchronize()

                def matvec(x, y, scale):
                    torch.cuda.current_stream(dv).synchronize()
                    L.check(f(f_data, _dptr(x), _dptr(y), float(scale)))
                    ctx.synchronize()

                kernel.compute(beta, matvec, a, b, [])
                torch.cuda.current_stream(dv).synchronize()
                L.check(lib.b2g_vec_copy(ctx.h, _dptr(b), C.c_void_p(b_ptr), n))
                ctx.synchronize()
                return 0
            except Exception:  # reported as B2G_ERR_STATE by the library
                import traceback
                traceback.print_exc()
                return 1

        self._kernel_cb = L.KERNEL_FN(trampoline)
        L.check(self.lib.b2g_effham_set_kernel(
            self.h, C.cast(self._kernel_cb, C.c_void_p), None))

    def set_davidson_mode(self, mode):
        """1: graph-captured device Davidson iteration (default, one root);
        0: host-driven loop."""
        L.check(self.lib.b2g_effham_set_davidson_mode(self.h, int(mode)))

    # ------------------------------------------------------------ backward
    def operator_data(self):
        """Copies of the left and right operator buffers (float64 device
        tensors; element offsets as in the problem's operator sets)."""
        lp, rp = C.c_void_p(), C.c_void_p()
        ln, rn = C.c_size_t(), C.c_size_t()
        L.check(self.lib.b2g_effham_data(self.h, C.byref(lp), C.byref(ln),
                                         C.byref(rp), C.byref(rn)))
        lt = torch.empty(ln.value, dtype=torch.float64, device=self.ctx.device)
        rt = torch.empty(rn.value, dtype=torch.float64, device=self.ctx.device)
        torch.cuda.current_stream(self.ctx.device).synchronize()
        for src, dst in ((lp, lt), (rp, rt)):
            if dst.numel():
                L.check(self.lib.b2g_vec_copy(self.ctx.h, src, _dptr(dst),
                                              dst.numel()))
        self.ctx.synchronize()
        return lt, rt

    def set_operator_data(self, left, right):
        """Replace the operator numbers (same structure): copies `left` and
        `right` in, refreshes the FP32 shadows and the diagonal."""
        lp, rp = C.c_void_p(), C.c_void_p()
        ln, rn = C.c_size_t(), C.c_size_t()
        L.check(self.lib.b2g_effham_data(self.h, C.byref(lp), C.byref(ln),
                                         C.byref(rp), C.byref(rn)))
        if left.numel() != ln.value or right.numel() != rn.value:
            raise ValueError("operator buffers hold %d / %d values" % (
                ln.value, rn.value))
        torch.cuda.current_stream(self.ctx.device).synchronize()
        L.check(self.lib.b2g_effham_set_data(self.h, _dptr(left),
                                             _dptr(right)))
        self._diag = None
        self.ctx.synchronize()

    def grad_ops(self, v, c):
        """Partial expectations: (dE/dL, dE/dR) of E = sum_r <v_r|H|c_r>
        with respect to every element of the left and right operator
        buffers (operator_functions.hpp:354-472 contract the same)."""
        nrhs = self._check(c, v)
        lp, rp = C.c_void_p(), C.c_void_p()
        ln, rn = C.c_size_t(), C.c_size_t()
        L.check(self.lib.b2g_effham_data(self.h, C.byref(lp), C.byref(ln),
                                         C.byref(rp), C.byref(rn)))
        gl = torch.empty(ln.value, dtype=torch.float64, device=self.ctx.device)
        gr = torch.empty(rn.value, dtype=torch.float64, device=self.ctx.device)
        torch.cuda.current_stream(self.ctx.device).synchronize()
        L.check(self.lib.b2g_effham_grad_ops(self.h, _dptr(v), _dptr(c), nrhs,
                                             _dptr(gl), _dptr(gr), 0))
        self.ctx.synchronize()
        return gl, gr

    def grad_flops(self):
        """Algorithmic FLOPs of grad_ops per right-hand side."""
        f = C.c_double()
        L.check(self.lib.b2g_grad_flops(self.plan, C.byref(f)))
        return f.value

    def deallocate(self):
        if self.h:
            self.lib.b2g_effham_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.deallocate()
        except Exception:
            pass


class Energy(torch.autograd.Function):
    """E(psi, L, R) = <psi| H(L, R) |psi> with a hand-written backward:
    dE/dpsi = 2 H psi (H_eff is real symmetric in DMRG), dE/dL and dE/dR
    from the partial expectations (EffectiveHamiltonian.grad_ops).
    Usage: Energy.apply(psi, left_data, right_data, heff)."""

    @staticmethod
    def forward(ctx, psi, left, right, heff):
        heff.set_operator_data(left.detach().contiguous(),
                               right.detach().contiguous())
        p = psi.detach().contiguous()
        sigma = heff.multiply(p)
        ctx.heff = heff
        ctx.save_for_backward(p, sigma)
        return torch.dot(p, sigma)

    @staticmethod
    def backward(ctx, g):
        p, sigma = ctx.saved_tensors
        gpsi = gl = gr = None
        if ctx.needs_input_grad[0]:
            gpsi = 2.0 * g * sigma
        if ctx.needs_input_grad[1] or ctx.needs_input_grad[2]:
            gl, gr = ctx.heff.grad_ops(p, p)
            gl, gr = g * gl, g * gr
        return gpsi, gl, gr, None


def connections(h, which):
    import numpy as np
    lib = L.lib()
    n = C.c_uint64()
    L.check(lib.b2g_effham_connections(h, which, C.byref(n), None, None, None,
                                       None, None))
    nc = n.value
    ia = np.zeros(nc, np.uint32)
    ib = np.zeros(nc, np.uint32)
    ic = np.zeros(nc, np.uint32)
    st = np.zeros(nc, np.uint64)
    f = np.zeros(nc, np.float64)
    L.check(lib.b2g_effham_connections(
        h, which, C.byref(n), L.ptr(ia, C.c_uint32), L.ptr(ib, C.c_uint32),
        L.ptr(ic, C.c_uint32), L.ptr(st, C.c_uint64), L.ptr(f, C.c_double)))
    return dict(ia=ia, ib=ib, ic=ic, stride=st, factor=f)


def host_plan(problem):
    """Build only the host-side plan (no device): returns a handle whose
    connection lists and pairs can be queried."""
    lib = L.lib()
    h = C.c_void_p()
    cp = problem.to_c()
    L.check(lib.b2g_effham_create(None, C.byref(cp), L.FLAG_HOST_ONLY,
                                  C.byref(h)))
    return h


def host_plan_info(h):
    lib = L.lib()
    kl, bl, npairs, macs = (C.c_uint64(), C.c_uint64(), C.c_uint64(),
                            C.c_uint64())
    L.check(lib.b2g_effham_info(h, C.byref(kl), C.byref(bl), C.byref(npairs),
                                C.byref(macs)))
    return dict(ket_len=kl.value, bra_len=bl.value, n_pairs=npairs.value,
                macs=macs.value)


def destroy(h):
    L.lib().b2g_effham_destroy(h)
