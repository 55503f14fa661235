"""Text checkpoint headers on the CPU: cntrx_herm_file_header and cntrx_function_file_header
read the first line the reference's print_to_file writes ("# nt ntau size1  sig",
libcntr/cntr/cntr_herm_matrix_impl.hpp:1313-1314; "# nt size1",
libcntr/cntr/cntr_function_impl.hpp:724) and reject a file without one."""
import ctypes
import os

import pytest

from nessi_amd import _lib


def _lib_or_skip():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libcntrx.so not built")
    return _lib.load()


def test_herm_header(tmp_path):
    lib = _lib_or_skip()
    f = tmp_path / "g.dat"
    f.write_text("# 400 300 2  -1\nmat: 0 1 0 0 0 0 0 1 0\n")
    v = [ctypes.c_int() for _ in range(4)]
    assert lib.cntrx_herm_file_header(os.fsencode(str(f)), *[ctypes.byref(x) for x in v]) == 0
    assert [x.value for x in v] == [400, 300, 2, -1]
    f.write_text("mat: 0 1 0\n")
    assert lib.cntrx_herm_file_header(os.fsencode(str(f)), *[ctypes.byref(x) for x in v]) != 0
    assert b"error in file" in lib.cntrx_last_error()
    assert lib.cntrx_herm_file_header(os.fsencode(str(tmp_path / "none")),
                                      *[ctypes.byref(x) for x in v]) != 0


def test_function_header(tmp_path):
    lib = _lib_or_skip()
    f = tmp_path / "f.dat"
    f.write_text("# 12 4\n1 0 \n")
    n, s = ctypes.c_int(), ctypes.c_int()
    assert lib.cntrx_function_file_header(os.fsencode(str(f)), ctypes.byref(n), ctypes.byref(s)) == 0
    assert (n.value, s.value) == (12, 4)
    f.write_text("")
    assert lib.cntrx_function_file_header(os.fsencode(str(f)), ctypes.byref(n), ctypes.byref(s)) != 0
    # shapes no print_to_file writes are rejected before anything is sized from them
    for bad in ("# -3 1\n", "# 4 0\n", "# 4 99999999999\n"):
        f.write_text(bad)
        assert lib.cntrx_function_file_header(os.fsencode(str(f)), ctypes.byref(n), ctypes.byref(s)) != 0


def test_herm_header_rejects_bad_shapes(tmp_path):
    lib = _lib_or_skip()
    f = tmp_path / "g.dat"
    v = [ctypes.c_int() for _ in range(4)]
    for bad in ("# -2 10 1  -1\n", "# 5 -1 1  -1\n", "# 5 10 0  -1\n"):
        f.write_text(bad)
        assert lib.cntrx_herm_file_header(os.fsencode(str(f)), *[ctypes.byref(x) for x in v]) != 0
    f.write_text("# -1 10 1  -1\n")  # Matsubara only
    assert lib.cntrx_herm_file_header(os.fsencode(str(f)), *[ctypes.byref(x) for x in v]) == 0


def test_header_of_a_directory_fails_with_a_status(tmp_path):
    """a path that opens but is not a regular file (a directory) is rejected with a status
    code; the reader never sizes a buffer from it and nothing escapes the C ABI"""
    lib = _lib_or_skip()
    v = [ctypes.c_int() for _ in range(4)]
    assert lib.cntrx_herm_file_header(os.fsencode(str(tmp_path)), *[ctypes.byref(x) for x in v]) != 0
    assert lib.cntrx_function_file_header(os.fsencode(str(tmp_path)), ctypes.byref(v[0]),
                                          ctypes.byref(v[1])) != 0
