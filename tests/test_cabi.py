"""CPU checks of the C-ABI boundary: libcntrx.so loads, exports every symbol that
include/cntrx.h declares, reports argument errors as status codes, and ships the same
Integrator weights as the (reference-pinned) oracle.  No GPU compute is invoked."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "cntrx.h")).read()
    return sorted(set(re.findall(r"\b(cntrx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from nessi_amd import _lib
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 18
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, name


@pytest.mark.parametrize("k", range(6))
def test_library_weights_equal_oracle(orc, k):
    from nessi_amd import cntr
    a = cntr.integrator_tables(k)
    b = orc.integrator_tables(k)
    for key in a:
        assert np.array_equal(a[key], b[key]), key


def test_invalid_arguments_return_codes():
    from nessi_amd import _lib
    lib = _lib.load()
    # NULL descriptors are rejected before any device work
    rc = lib.cntrx_convolution_timestep(3, None, None, None, None, None, 1.0, 0.1, 5, None)
    assert rc in (_lib.CNTRX_EINVAL, _lib.CNTRX_EHIP)
    assert lib.cntrx_integrator_tables(9, None, None, None, None, None, None, None) == _lib.CNTRX_EINVAL
    assert b"range" in lib.cntrx_last_error()


def test_noalloc_rejects_a_timestep_beyond_its_scratch():
    """cntrx_dyson_timestep_noalloc checks its fixed scratch size before it touches the
    device: size1 8 at nt 6000 needs (3*6001+8)*64*16 bytes of workspace, more than 16 MiB.
    The descriptors' pointers are never dereferenced on this path."""
    from nessi_amd import _lib
    lib = _lib.load()
    fake = ctypes.c_void_p(0x1000)
    G = _lib.CntrxHerm(6000, 10, 8, -1, fake, fake, fake, fake)
    S = _lib.CntrxHerm(6000, 10, 8, -1, fake, fake, fake, fake)
    H = _lib.CntrxFunc(6000, 8, fake)
    rc = lib.cntrx_dyson_timestep_noalloc(6000, ctypes.byref(G), 0.0, ctypes.byref(H), ctypes.byref(S),
                                          1.0, 0.01, 5, None)
    assert rc == _lib.CNTRX_EINVAL
    msg = lib.cntrx_last_error()
    assert b"scratch" in msg and b"16777216" in msg, msg


def test_herm_matrix_window_rejects_a_window_too_short_for_the_predictor():
    """the predictor reads the k+1 timesteps before the one it fills: a window must hold
    at least 2 (k + 2) timesteps (checked before any device memory is touched)"""
    from nessi_amd import cntr
    with pytest.raises(ValueError):
        cntr.herm_matrix_window(50, 10, 1, window=13, SolveOrder=5)


def test_green_from_H_td_rejects_bad_arguments_before_touching_a_device():
    """cf_order other than 2/4, a timestep beyond G, H shorter than the propagation needs:
    EINVAL with a message, checked before the device is initialised"""
    import ctypes
    from nessi_amd import _lib
    lib = _lib.load()
    nt, ntau, s = 10, 4, 2
    mat = (ctypes.c_double * (2 * (ntau + 1) * s * s))()
    big = (ctypes.c_double * (2 * 200 * s * s))()
    G = _lib.CntrxHerm(nt, ntau, s, -1, ctypes.addressof(big), ctypes.addressof(big), ctypes.addressof(big),
                       ctypes.addressof(mat))
    hdat = (ctypes.c_double * (2 * (nt + 2) * s * s))()
    H = _lib.CntrxFunc(nt, s, ctypes.addressof(hdat))
    Hshort = _lib.CntrxFunc(3, s, ctypes.addressof(hdat))
    for args in ((H, 3, -2), (H, 4, nt + 1), (Hshort, 4, -2)):
        f, cf, tstp = args
        rc = lib.cntrx_green_from_H_td(ctypes.byref(G), 0.0, ctypes.byref(f), 2.0, 0.1, 5, cf, 1, tstp, None)
        assert rc == _lib.CNTRX_EINVAL, args
        assert b"green_from_H" in lib.cntrx_last_error()
