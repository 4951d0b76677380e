"""block2's sparse-matrix file format (b2g_smat_*), on the CPU.

The golden fixtures hold the Davidson vector of each dumped site as block2
wrote it (SparseMatrix::save_data with save_info, oracle/ref_harness.cpp)
next to the same vector and its info as arrays.  Checked: the reader
returns exactly those arrays (structure, factor, data), and the writer
reproduces the reference's file byte for byte; a truncated file and
compressed storage are refused.
"""
import numpy as np
import pytest

from conftest import FIXTURES, DELAYED_FIXTURES, fixture_path

from b2g import formats as F
from b2g import _lib as L


@pytest.mark.parametrize("name", FIXTURES + DELAYED_FIXTURES)
def test_read_matches_reference_arrays(tmp_path, name):
    z = np.load(fixture_path(name))
    if "ket_dav_smat" not in z.files:
        pytest.skip("fixture without the file dump")
    p = tmp_path / "ket.smat"
    p.write_bytes(z["ket_dav_smat"].tobytes())
    info, factor, data, flags = F.load_sparse_matrix(p)
    assert factor == 1.0
    assert np.array_equal(data, z["psi_dav"])
    assert np.array_equal(info.quanta, z["ket_quanta"].astype(np.uint64))
    assert np.array_equal(info.nbra, z["ket_nbra"])
    assert np.array_equal(info.nket, z["ket_nket"])
    assert np.array_equal(info.shift, z["ket_shift"].astype(np.uint64))
    assert info.dq == int(z["ket_dq"][0])
    assert info.fermion == int(z["ket_fermion"][0])
    assert flags & 1  # a wavefunction
    # the writer reproduces the reference's bytes
    q = tmp_path / "out.smat"
    F.save_sparse_matrix(q, data, info, factor, flags)
    assert q.read_bytes() == p.read_bytes()
    # the matrix part alone (save_data without save_info)
    r = tmp_path / "data.smat"
    F.save_sparse_matrix(r, data, None, 0.5)
    none, f2, d2, _ = F.load_sparse_matrix(r, with_info=False)
    assert none is None and f2 == 0.5 and np.array_equal(d2, data)


def test_refuses_bad_files(tmp_path):
    z = np.load(fixture_path("hub8_su2"))
    raw = z["ket_dav_smat"].tobytes()
    p = tmp_path / "cut.smat"
    p.write_bytes(raw[:len(raw) - 8])
    with pytest.raises(L.B2GError):
        F.load_sparse_matrix(p)
    # compressed storage marker in place of the element count
    info, factor, data, flags = F.load_sparse_matrix(_write(tmp_path, raw))
    ninfo = 8 + 4 + 5 * 4 * len(info.quanta) + 2
    bad = raw[:ninfo + 8] + (2 ** 64 - 1).to_bytes(8, "little") + \
        raw[ninfo + 16:]
    with pytest.raises(L.B2GError):
        F.load_sparse_matrix(_write(tmp_path, bad))


def _write(tmp_path, b):
    p = tmp_path / "x.smat"
    p.write_bytes(b)
    return p
