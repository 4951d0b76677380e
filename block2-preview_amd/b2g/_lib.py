"""ctypes binding of libb2g.so (the C ABI declared in include/b2g.h).

The library is the product: there is no Python or CPU fallback.  Loading
fails loudly when the shared object has not been built.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libb2g.so")

B2G_OK = 0
B2G_ERR_NOCONV = 4
B2G_F64, B2G_F32 = 0, 1
PREC = {"f64": 0, "f32": 1, "bf16x3": 2}
B2G_SZ, B2G_SU2 = 0, 1
FLAG_F32 = 1
FLAG_BF16 = 2
FLAG_HOST_ONLY = 0x100
FLAG_DEVICE_DATA = 0x200

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
f64p = C.POINTER(C.c_double)


class b2g_pair(C.Structure):
    _fields_ = [("a_off", C.c_uint64), ("b_off", C.c_uint64),
                ("c_off", C.c_uint64), ("v_off", C.c_uint64),
                ("c_m", C.c_uint32), ("c_n", C.c_uint32),
                ("v_m", C.c_uint32), ("v_n", C.c_uint32),
                ("conj", C.c_uint32), ("kind", C.c_uint32),
                ("c_ld", C.c_uint32), ("v_ld", C.c_uint32),
                ("alpha", C.c_double)]


class b2g_block(C.Structure):
    _fields_ = [("off", C.c_uint64), ("m", C.c_uint32), ("n", C.c_uint32)]


class b2g_diag_term(C.Structure):
    _fields_ = [("a_off", C.c_uint64), ("b_off", C.c_uint64),
                ("d_off", C.c_uint64), ("a_n", C.c_uint32),
                ("b_n", C.c_uint32), ("r0", C.c_uint32), ("c0", C.c_uint32),
                ("alpha", C.c_double)]


class b2g_sparse_info(C.Structure):
    _fields_ = [("delta_quantum", C.c_uint64), ("is_fermion", C.c_int32),
                ("n", C.c_int32), ("quanta", u64p), ("n_states_bra", u32p),
                ("n_states_ket", u32p), ("block_shifts", u64p)]


class b2g_operator(C.Structure):
    _fields_ = [("delta_quantum", C.c_uint64), ("factor", C.c_double),
                ("info", C.c_int32), ("flags", C.c_int32),
                ("data_off", C.c_uint64)]


class b2g_term(C.Structure):
    _fields_ = [("a", C.c_int32), ("b", C.c_int32), ("conj", C.c_uint32),
                ("kind", C.c_uint32), ("factor", C.c_double),
                ("da", C.c_int32), ("db", C.c_int32), ("dconj", C.c_uint32),
                ("pad", C.c_uint32)]


class b2g_tp_conn(C.Structure):
    _fields_ = [("n", C.c_int32 * 5), ("nc", C.c_int32), ("quanta", u64p),
                ("idx", u32p), ("ia", u32p), ("ib", u32p), ("ic", u32p),
                ("stride", u64p), ("factor", f64p)]


class b2g_problem(C.Structure):
    _fields_ = [("symmetry", C.c_int32), ("n_left_infos", C.c_int32),
                ("n_right_infos", C.c_int32), ("n_left_ops", C.c_int32),
                ("n_right_ops", C.c_int32), ("n_terms", C.c_int32),
                ("opdq", C.c_uint64), ("hop_left_vacuum", C.c_uint64),
                ("ket", b2g_sparse_info), ("bra", b2g_sparse_info),
                ("left_info_dq", u64p), ("right_info_dq", u64p),
                ("left_infos", C.POINTER(b2g_sparse_info)),
                ("right_infos", C.POINTER(b2g_sparse_info)),
                ("left_ops", C.POINTER(b2g_operator)),
                ("right_ops", C.POINTER(b2g_operator)),
                ("terms", C.POINTER(b2g_term)),
                ("left_data", f64p), ("right_data", f64p),
                ("left_len", C.c_size_t), ("right_len", C.c_size_t),
                ("delayed_side", C.c_int32), ("n_dl_infos", C.c_int32),
                ("n_dr_infos", C.c_int32), ("n_dl_ops", C.c_int32),
                ("n_dr_ops", C.c_int32), ("pad_delayed", C.c_int32),
                ("dl_infos", C.POINTER(b2g_sparse_info)),
                ("dr_infos", C.POINTER(b2g_sparse_info)),
                ("dl_ops", C.POINTER(b2g_operator)),
                ("dr_ops", C.POINTER(b2g_operator)),
                ("tp", C.POINTER(b2g_tp_conn))]


class b2g_davidson_opts(C.Structure):
    _fields_ = [("conv_thrd", C.c_double), ("rel_conv_thrd", C.c_double),
                ("shift", C.c_double), ("max_iter", C.c_int),
                ("soft_max_iter", C.c_int), ("def_min", C.c_int),
                ("def_max", C.c_int), ("davidson_type", C.c_int),
                ("precision", C.c_int), ("refine", C.c_int),
                ("n_ortho", C.c_int), ("ortho", C.POINTER(C.c_void_p)),
                ("proj_weights", f64p)]


class b2g_state_info(C.Structure):
    _fields_ = [("n", C.c_int32), ("pad", C.c_int32), ("quanta", u64p),
                ("n_states", u32p)]


class b2g_contract_term(C.Structure):
    _fields_ = [("c", C.c_int32), ("a", C.c_int32), ("b", C.c_int32),
                ("conj", C.c_uint32), ("factor", C.c_double)]


class b2g_transform_term(C.Structure):
    _fields_ = [("c", C.c_int32), ("a", C.c_int32), ("conj", C.c_uint32),
                ("pad", C.c_uint32), ("factor", C.c_double)]


class b2g_contract(C.Structure):
    _fields_ = [("symmetry", C.c_int32), ("n_terms", C.c_int32),
                ("bra_a", b2g_state_info), ("ket_a", b2g_state_info),
                ("bra_b", b2g_state_info), ("ket_b", b2g_state_info),
                ("bra", b2g_state_info), ("ket", b2g_state_info),
                ("n_a_infos", C.c_int32), ("n_b_infos", C.c_int32),
                ("n_c_infos", C.c_int32), ("n_a_ops", C.c_int32),
                ("n_b_ops", C.c_int32), ("n_c_ops", C.c_int32),
                ("a_info_dq", u64p), ("b_info_dq", u64p), ("c_info_dq", u64p),
                ("a_infos", C.POINTER(b2g_sparse_info)),
                ("b_infos", C.POINTER(b2g_sparse_info)),
                ("c_infos", C.POINTER(b2g_sparse_info)),
                ("a_ops", C.POINTER(b2g_operator)),
                ("b_ops", C.POINTER(b2g_operator)),
                ("c_ops", C.POINTER(b2g_operator)),
                ("terms", C.POINTER(b2g_contract_term))]


# DavidsonTypes bits taken by b2g_effham_eigs_ex (b2g.h B2G_DAV_*)
DAV = {"Normal": 0, "GreaterThan": 1, "LessThan": 2, "CloseTo": 4,
       "DavidsonPrecond": 32, "NoPrecond": 64, "Exact": 256,
       "ElementProj": 1024}


# every symbol declared in include/b2g.h, with (restype, argtypes)
_vp = C.c_void_p

# matvec hook (b2g.h b2g_matvec_fn / b2g_kernel_fn)
MATVEC_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                        C.c_double)
KERNEL_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, MATVEC_FN,
                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t)

# kernel ids of one H*psi (b2g.h): phase-1 / phase-2 grouped GEMM per tile
# configuration c = 3*i + j (TM = 32 << i, TN = 32 << j), split-K reduction
K_PHASE1, K_PHASE2, K_SPLITK_REDUCE, K_PACK, NKERNELS = 0, 9, 18, 19, 20


def kernel_name(k):
    if k == K_SPLITK_REDUCE:
        return "splitk_reduce"
    if k == K_PACK:
        return "pack_bf16"
    c = k % 9
    return "phase%d_%dx%d" % (1 if k < K_PHASE2 else 2, 32 << (c // 3),
                              32 << (c % 3))

_SIGS = {
    "b2g_last_error": (C.c_char_p, []),
    "b2g_abi_sizes": (C.c_int, [C.POINTER(C.c_size_t), C.c_int]),
    "b2g_ctx_create": (C.c_int, [C.c_int, C.c_uint, C.POINTER(_vp)]),
    "b2g_ctx_destroy": (C.c_int, [_vp]),
    "b2g_ctx_synchronize": (C.c_int, [_vp]),
    "b2g_ctx_stream": (_vp, [_vp]),
    "b2g_plan_set_timing": (C.c_int, [_vp, C.c_int]),
    "b2g_plan_timing": (C.c_int, [_vp, f64p, u64p]),
    "b2g_plan_kernel_flops": (C.c_int, [_vp, f64p]),
    "b2g_plan_kernel_flops_for": (C.c_int, [_vp, C.c_int, f64p]),
    "b2g_effham_plan": (_vp, [_vp]),
    "b2g_set_tile_rates": (C.c_int, [f64p, C.c_int]),
    "b2g_set_tile_rates_for": (C.c_int, [C.c_int, f64p, C.c_int]),
    "b2g_get_tile_rates_for": (C.c_int, [C.c_int, f64p, C.c_int]),
    "b2g_ctx_num_cu": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "b2g_buf_create": (C.c_int, [_vp, C.c_size_t, C.c_int, C.POINTER(_vp)]),
    "b2g_buf_destroy": (C.c_int, [_vp]),
    "b2g_upload_blocks": (C.c_int, [_vp, _vp, C.c_size_t, C.c_int,
                                    C.POINTER(_vp)]),
    "b2g_buf_write": (C.c_int, [_vp, _vp, C.c_size_t, C.c_size_t]),
    "b2g_buf_read": (C.c_int, [_vp, _vp, C.c_size_t, C.c_size_t]),
    "b2g_buf_device_ptr": (_vp, [_vp]),
    "b2g_buf_size": (C.c_size_t, [_vp]),
    "b2g_upload_plan": (C.c_int, [_vp, _vp, _vp, C.POINTER(b2g_pair),
                                  C.c_size_t, C.POINTER(b2g_block),
                                  C.c_size_t, C.c_size_t, C.c_size_t,
                                  C.c_uint, C.POINTER(_vp)]),
    "b2g_plan_destroy": (C.c_int, [_vp]),
    "b2g_plan_macs": (C.c_int, [_vp, u64p]),
    "b2g_plan_stats": (C.c_int, [_vp, u64p, u64p, u64p]),
    "b2g_hpsi": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_double, C.c_int,
                           C.c_int]),
    "b2g_effham_diag_stats": (C.c_int, [_vp, u64p, u64p, u64p]),
    "b2g_hpsi_noalloc": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_double,
                                   C.c_int, C.c_int]),
    "b2g_diag": (C.c_int, [_vp, _vp, _vp, C.POINTER(b2g_diag_term),
                           C.c_size_t, C.POINTER(b2g_block), C.c_size_t, _vp,
                           C.c_size_t]),
    "b2g_vec_dot": (C.c_int, [_vp, _vp, _vp, C.c_size_t, f64p]),
    "b2g_vec_axpy": (C.c_int, [_vp, C.c_double, _vp, _vp, C.c_size_t]),
    "b2g_vec_scale": (C.c_int, [_vp, C.c_double, _vp, C.c_size_t]),
    "b2g_vec_copy": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "b2g_effham_create": (C.c_int, [_vp, C.POINTER(b2g_problem), C.c_uint,
                                    C.POINTER(_vp)]),
    "b2g_effham_destroy": (C.c_int, [_vp]),
    "b2g_effham_info": (C.c_int, [_vp, u64p, u64p, u64p, u64p]),
    "b2g_effham_connections": (C.c_int, [_vp, C.c_int, u64p, u32p, u32p,
                                         u32p, u64p, f64p]),
    "b2g_effham_pairs": (C.c_int, [_vp, u64p, C.POINTER(b2g_pair)]),
    "b2g_effham_hpsi": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_double,
                                  C.c_int, C.c_int]),
    "b2g_effham_diag": (C.c_int, [_vp, _vp]),
    "b2g_effham_set_data": (C.c_int, [_vp, _vp, _vp]),
    "b2g_effham_data": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(C.c_size_t),
                                  C.POINTER(_vp), C.POINTER(C.c_size_t)]),
    "b2g_effham_grad_ops": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp, _vp,
                                      C.c_int]),
    "b2g_grad_ops": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp, _vp,
                               C.c_int]),
    "b2g_grad_flops": (C.c_int, [_vp, f64p]),
    "b2g_plan_refresh": (C.c_int, [_vp]),
    "b2g_plan_reserve": (C.c_int, [_vp, C.c_int, C.c_int]),
    "b2g_effham_set_davidson_mode": (C.c_int, [_vp, C.c_int]),
    "b2g_effham_set_kernel": (C.c_int, [_vp, _vp, _vp]),
    "b2g_effham_eigs": (C.c_int, [_vp, _vp, C.c_int, C.c_double, C.c_int,
                                  C.c_int, C.c_int, C.c_int, C.c_int, f64p,
                                  C.POINTER(C.c_int), u64p, f64p]),
    "b2g_davidson_opts_default": (None, [C.POINTER(b2g_davidson_opts)]),
    "b2g_effham_eigs_ex": (C.c_int, [_vp, _vp, C.c_int,
                                     C.POINTER(b2g_davidson_opts), f64p,
                                     C.POINTER(C.c_int), u64p, f64p]),
    "b2g_syev_batched": (C.c_int, [_vp, C.c_int, u32p, u64p, _vp, _vp, _vp,
                                   C.POINTER(C.c_int)]),
    "b2g_split_density_matrix": (C.c_int, [_vp, C.c_int,
                                           C.POINTER(b2g_sparse_info), _vp,
                                           C.c_int, C.c_double, C.c_int,
                                           C.c_int, C.c_uint, C.c_double,
                                           C.POINTER(_vp)]),
    "b2g_split_result": (C.c_int, [_vp, C.c_int, C.POINTER(b2g_sparse_info),
                                   C.POINTER(_vp), C.POINTER(C.c_size_t)]),
    "b2g_split_error": (C.c_int, [_vp, f64p]),
    "b2g_split_values": (C.c_int, [_vp, C.c_int, f64p,
                                   C.POINTER(C.c_size_t)]),
    "b2g_split_destroy": (C.c_int, [_vp]),
    "b2g_split_density_matrix_multi": (C.c_int, [
        _vp, C.c_int, C.POINTER(b2g_sparse_info), _vp, C.c_int, f64p,
        C.c_int, C.c_double, C.c_int, C.c_int, C.c_uint, C.c_double,
        C.POINTER(_vp)]),
    "b2g_split_density_matrix_extra": (C.c_int, [
        _vp, C.c_int, C.POINTER(b2g_sparse_info), _vp, C.c_int,
        C.POINTER(b2g_sparse_info), u64p, C.c_int, C.c_double, C.c_int,
        C.c_int, C.c_uint, C.c_double, C.POINTER(_vp)]),
    "b2g_split_n_roots": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "b2g_split_density_matrix_noise": (C.c_int, [
        _vp, C.c_int, C.POINTER(b2g_sparse_info), _vp, C.c_int, f64p,
        C.c_double, C.c_int, C.c_uint, C.c_int, C.c_double, C.c_int,
        C.c_int, C.c_uint, C.c_double, C.POINTER(_vp)]),
    "b2g_smat_load": (C.c_int, [C.c_char_p, C.c_int, C.c_int,
                                C.POINTER(_vp)]),
    "b2g_smat_get": (C.c_int, [_vp, C.POINTER(b2g_sparse_info),
                               C.POINTER(C.c_int), f64p, C.POINTER(f64p),
                               C.POINTER(C.c_size_t)]),
    "b2g_smat_save": (C.c_int, [C.c_char_p, C.POINTER(b2g_sparse_info),
                                C.c_int, C.c_int, C.c_double, f64p,
                                C.c_size_t]),
    "b2g_smat_destroy": (C.c_int, [_vp]),
    "b2g_contract_plan": (C.c_int, [_vp, C.POINTER(b2g_contract), C.c_uint,
                                    C.POINTER(_vp)]),
    "b2g_winplan_connections": (C.c_int, [
        _vp, C.c_int, C.POINTER(C.c_int32), u64p, u64p, u32p, u64p, u32p,
        u32p, u32p, u64p, f64p]),
    "b2g_transform_plan": (C.c_int, [
        _vp, C.c_int, C.c_int, u64p, C.POINTER(b2g_sparse_info), C.c_int,
        C.POINTER(b2g_operator), C.c_int, C.POINTER(b2g_transform_term),
        C.c_uint, C.POINTER(_vp)]),
    "b2g_swap_plan": (C.c_int, [
        _vp, C.c_int, C.c_int, C.POINTER(b2g_state_info),
        C.POINTER(b2g_state_info), C.POINTER(b2g_state_info),
        C.POINTER(b2g_state_info), C.POINTER(b2g_state_info),
        C.POINTER(b2g_sparse_info), C.POINTER(b2g_sparse_info), C.c_uint,
        C.POINTER(_vp)]),
    "b2g_winplan_run": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int]),
    "b2g_winplan_stats": (C.c_int, [_vp, u64p, u64p, u64p, u64p, u64p]),
    "b2g_winplan_destroy": (C.c_int, [_vp]),
    "b2g_rotate_plan": (C.c_int, [
        _vp, C.c_int, C.c_int, C.POINTER(b2g_sparse_info), _vp, C.c_int,
        u64p, C.POINTER(b2g_sparse_info), C.c_int, u64p,
        C.POINTER(b2g_sparse_info), C.c_int, C.POINTER(b2g_operator),
        C.POINTER(b2g_operator), C.c_size_t, C.c_size_t, C.c_uint,
        C.POINTER(_vp)]),
    "b2g_contract_wfn_plan": (C.c_int, [
        _vp, C.c_int, C.POINTER(b2g_sparse_info), C.c_int,
        C.POINTER(b2g_sparse_info), C.POINTER(b2g_sparse_info), C.c_double,
        C.POINTER(_vp)]),
    "b2g_gemmlist_run": (C.c_int, [_vp, _vp, _vp, _vp]),
    "b2g_gemmlist_stats": (C.c_int, [_vp, u64p, u64p]),
    "b2g_gemmlist_destroy": (C.c_int, [_vp]),
    "b2g_rng_create": (C.c_int, [C.c_uint, C.POINTER(_vp)]),
    "b2g_rng_destroy": (C.c_int, [_vp]),
    "b2g_split_density_matrix_noise_rng": (C.c_int, [
        _vp, C.c_int, C.POINTER(b2g_sparse_info), _vp, C.c_int, f64p,
        C.c_double, C.c_int, _vp, C.c_int, C.c_double, C.c_int,
        C.c_int, C.c_uint, C.c_double, C.POINTER(_vp)]),
    "b2g_buf_wrap": (C.c_int, [_vp, _vp, C.c_size_t, C.c_int,
                               C.POINTER(_vp)]),
}

_lib = None


class B2GError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("b2g error %d: %s" % (code, msg))
        self.code = code


def lib():
    """Load libb2g.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libb2g.so not found at %s: build it first "
                "(python -c 'import __graft_entry__ as g; g.build()')" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != B2G_OK:
        msg = lib().b2g_last_error()
        raise B2GError(code, msg.decode() if msg else "")
    return code


def ptr(a, ctype):
    """ctypes pointer to a contiguous numpy array."""
    return a.ctypes.data_as(C.POINTER(ctype))


ABI_STRUCTS = ["b2g_pair", "b2g_block", "b2g_diag_term", "b2g_sparse_info",
               "b2g_operator", "b2g_term", "b2g_tp_conn", "b2g_problem",
               "b2g_davidson_opts"]
