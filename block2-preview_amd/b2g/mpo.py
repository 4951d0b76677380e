"""block2's MPO and MPS files, read for the sweep driver.

Serialisations followed:
  MPO::save_data                    src/dmrg/mpo.hpp:795-886
  MPSInfo::save_data                src/dmrg/mps.hpp:169-191
  StateInfo::save_data              src/core/state_info.hpp:101-105
  SparseMatrixInfo::save_data       src/core/sparse_matrix.hpp:555-575
  SparseMatrix::save_data(file, save_info)
                                    src/core/sparse_matrix.hpp:954-989
  OperatorTensor::save_data         src/core/operator_tensor.hpp:122-148
  OpExpr / Symbolic                 src/core/expr.hpp, src/core/symbolic.hpp

`ubond` is the byte width of block2's ubond_t in the build that wrote the
files: 4 with _LARGE_BOND (total_bond_t is then 64-bit), 2 by default, 1
with _SMALL_BOND.  Only uncompressed storage is read.
"""
import gzip
import os

import numpy as np

from .effprob import FormatError, _Reader


def _state_info(r):
    """StateInfo::save_data: n_states_total, n, quanta, n_states."""
    if r.ubond == 4:
        r.take(8)
    else:
        r.take(4)
    n = r.i32()
    if n < 0 or n >= (1 << 28):
        raise FormatError("bad state count %d" % n)
    q = r.arr(np.uint64, n)
    words = {4: n, 2: n - (n >> 1), 1: (n + 3) >> 2}[r.ubond]
    raw = r.take(4 * words)
    ut = {4: np.uint32, 2: np.uint16, 1: np.uint8}[r.ubond]
    ns = np.frombuffer(raw, dtype=ut)[:n].astype(np.uint32)
    return dict(quanta=q, n_states=ns)


def _read_file(path):
    """The bytes of `path`, or of its gzip-compressed copy `path`.gz when
    only that exists."""
    if not os.path.exists(path) and os.path.exists(path + ".gz"):
        with gzip.open(path + ".gz", "rb") as f:
            return f.read()
    with open(path, "rb") as f:
        return f.read()


def load_mpo(path, ubond=4):
    """An MPO written by MPO::save_data (uncompressed operators).
    Returns a dict: n_sites, tag, const_e, left_vacuum, sparse_form, op (the
    Hamiltonian's OpElement), schemer (None or the transformation's sites,
    new operator names and expressions), site_op_infos[i] = [(dq, info)], tensors[i]
    (lmat / rmat symbolics and the site operators with their data), basis[i]
    (StateInfo), left/right/middle operator names and exprs (lists of
    symbolics, empty where the MPO has none)."""
    r = _Reader(_read_file(path), ubond)
    m = {}
    n = m["n_sites"] = r.i32()
    m["tag"] = bytes(r.take(r.i32())).decode()
    # const_e is a long double (x87 80-bit extended in 16 bytes on x86-64)
    m["const_e"] = float(np.frombuffer(bytes(r.take(16)),
                                       dtype=np.longdouble)[0])
    m["left_vacuum"] = r.u64()
    m["sparse_form"] = bytes(r.take(n)).decode()
    has_op, has_schemer = r.u8(), r.u8()
    m["op"] = r.expr() if has_op else None
    # MPOSchemer::save_data (src/dmrg/mpo.hpp:88-95): the operator
    # transformation at the switch between normal and complementary
    # operators
    m["schemer"] = None
    if has_schemer:
        lt, rt = r.u16(), r.u16()
        m["schemer"] = dict(
            left_trans_site=lt, right_trans_site=rt,
            left_new_operator_names=r.symbolic(),
            right_new_operator_names=r.symbolic(),
            left_new_operator_exprs=r.symbolic(),
            right_new_operator_exprs=r.symbolic())
    m["site_op_infos"] = []
    for _ in range(r.i32()):
        lst = []
        for _ in range(r.i32()):
            dq = r.u64()
            lst.append((dq, r.info()))
        m["site_op_infos"].append(lst)
    m["tensors"] = [r.data_tensor() for _ in range(r.i32())]
    m["basis"] = [_state_info(r) for _ in range(r.i32())]
    for key in ("left_operator_names", "right_operator_names",
                "middle_operator_names", "left_operator_exprs",
                "right_operator_exprs", "middle_operator_exprs"):
        m[key] = [r.symbolic() for _ in range(r.i32())]
    if r.p != len(r.b):
        raise FormatError("%d trailing bytes after the MPO" % (len(r.b) - r.p))
    return m


def load_mps_info(path, ubond=4):
    """MPSInfo::save_data: n_sites, vacuum, target, bond_dim, tag, basis,
    left_dims_fci[0..n], right_dims_fci[0..n]."""
    with open(path, "rb") as f:
        r = _Reader(f.read(), ubond)
    n = r.i32()
    info = dict(n_sites=n, vacuum=r.u64(), target=r.u64())
    info["bond_dim"] = {4: r.u32, 2: r.u16, 1: r.u8}[ubond]()
    info["tag"] = bytes(r.take(r.i32())).decode()
    info["basis"] = [_state_info(r) for _ in range(n)]
    info["left_dims_fci"] = [_state_info(r) for _ in range(n + 1)]
    info["right_dims_fci"] = [_state_info(r) for _ in range(n + 1)]
    return info


def load_sparse_matrix(path, ubond=4):
    """SparseMatrix::save_data(filename, save_info=true): the info, then
    factor, element count and elements."""
    with open(path, "rb") as f:
        r = _Reader(f.read(), ubond)
    info = r.info()
    factor = r.f64()
    total = r.u64()
    if total == (1 << 64) - 1:
        raise FormatError("compressed sparse-matrix storage is not supported")
    data = r.arr(np.float64, total)
    return dict(info=info, factor=factor, data=data,
                wavefunction=bool(info["flags"] & 1))


def load_mps(directory, ubond=4):
    """The MPS a sweep starts from: mps_info.bin, mps_<i>.bin for every
    site holding a tensor, and mps.txt (n_sites, center, dot, canonical
    form), as the reference harness writes them with block2's save_data."""
    with open(os.path.join(directory, "mps.txt")) as f:
        parts = f.read().split()
    n, center, dot = int(parts[0]), int(parts[1]), int(parts[2])
    form = parts[3] if len(parts) > 3 else ""
    info = load_mps_info(os.path.join(directory, "mps_info.bin"), ubond)
    tensors = []
    for i in range(n):
        p = os.path.join(directory, "mps_%d.bin" % i)
        tensors.append(load_sparse_matrix(p, ubond) if os.path.exists(p)
                       else None)
    return dict(info=info, tensors=tensors, center=center, dot=dot,
                canonical_form=form, n_sites=n)
