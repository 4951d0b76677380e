"""block2's binary sparse-matrix files (SparseMatrix::save_data /
load_data with the SparseMatrixInfo first, src/core/sparse_matrix.hpp:
504-575, 908-989), through libb2g's b2g_smat_* reader and writer."""
import ctypes as C

import numpy as np

from . import _lib as L
from .truncate import BlockInfo


def load_sparse_matrix(path, with_info=True, ubond_size=4):
    """Returns (info, factor, data, flags): info a BlockInfo (None without
    with_info), data a float64 numpy array, flags the info's last byte
    (block_shifts_fl_size << 2 | is_wavefunction)."""
    lib = L.lib()
    h = C.c_void_p()
    L.check(lib.b2g_smat_load(str(path).encode(), int(bool(with_info)),
                              int(ubond_size), C.byref(h)))
    try:
        ci = L.b2g_sparse_info()
        flags, factor = C.c_int(0), C.c_double(0)
        dp = C.POINTER(C.c_double)()
        n = C.c_size_t(0)
        L.check(lib.b2g_smat_get(h, C.byref(ci) if with_info else None,
                                 C.byref(flags), C.byref(factor), C.byref(dp),
                                 C.byref(n)))
        data = (np.ctypeslib.as_array(dp, (n.value,)).copy() if n.value
                else np.zeros(0))
        info = BlockInfo.from_c(ci) if with_info else None
    finally:
        lib.b2g_smat_destroy(h)
    return info, factor.value, data, flags.value


def save_sparse_matrix(path, data, info=None, factor=1.0, flags=(8 << 2) | 1,
                       ubond_size=4):
    """Write `data` (float64) as block2 does; with `info` (BlockInfo) the
    info comes first, as SparseMatrix::save_data(filename, true)."""
    data = np.ascontiguousarray(data, dtype=np.float64)
    ci = info.to_c() if info is not None else None
    L.check(L.lib().b2g_smat_save(str(path).encode(),
                                  C.byref(ci) if ci is not None else None,
                                  int(flags), int(ubond_size), float(factor),
                                  L.ptr(data, C.c_double), data.size))
