"""cntrx_dyson_timestep_noalloc (scratch in a fixed device array, no allocation) against
cntrx_dyson_timestep on the same inputs, through the C ABI: bit for bit the same G, on two
calls in a row, the second with a smaller problem (so it runs on what the first call left
in the shared scratch), for size1 1 and 8, then 2 and 4; and the oracle agrees to 1e-12."""
import ctypes

import numpy as np
import pytest

from tests.gpu_util import component_errors, func_to_gpu, to_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cntr():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nessi_amd import cntr as c
    from nessi_amd import _lib
    _lib.check(_lib.load().cntrx_init(0))
    return c


def _rand_herm(rng, s, scale=0.5):
    X = rng.normal(scale=scale, size=(s, s)) + 1j * rng.normal(scale=scale, size=(s, s))
    return 0.5 * (X + X.conj().T)


def _state(orc, s, nt, ntau, n, seed):
    """oracle solve up to timestep n - 1 and the oracle's timestep n (H, Sigma, X_before, X_after)"""
    rng = np.random.default_rng(seed)
    beta, h, k = 3.0, 0.04, 5
    H = orc.HostFunc(nt, s)
    H0, H1 = _rand_herm(rng, s), _rand_herm(rng, s, 0.2)
    H[-1] = H0
    for t in range(nt + 1):
        H[t] = H0 + np.sin(0.6 * t * h) * H1
    Sig = orc.HostHerm(nt, ntau, s)
    orc.green_from_H_const(Sig, 0.0, _rand_herm(rng, s), beta, h)
    for c in ("ret", "les", "tv", "mat"):
        getattr(Sig, c)[...] *= 0.16
    X = orc.HostHerm(nt, ntau, s)
    orc.dyson_mat(X, 0.1, H, Sig, beta, k, 2)
    orc.dyson_start(X, 0.1, H, Sig, beta, h, k)
    for t in range(k + 1, n):
        orc.dyson_timestep(t, X, 0.1, H, Sig, beta, h, k)
    before = X.copy()
    orc.dyson_timestep(n, X, 0.1, H, Sig, beta, h, k)
    return H, Sig, before, X, beta, h, k


def _one(cntr, orc, s, nt, ntau, n, seed):
    import torch
    from nessi_amd import _lib
    lib = _lib.load()
    H, Sig, before, after, beta, h, k = _state(orc, s, nt, ntau, n, seed)
    Hg, Sg = func_to_gpu(cntr, H), to_gpu(cntr, Sig)
    Ga, Gb = to_gpu(cntr, before), to_gpu(cntr, before)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    print("no-alloc case: size1=%d nt=%d ntau=%d n=%d" % (s, nt, ntau, n), flush=True)
    _lib.check(lib.cntrx_dyson_timestep(n, Ga.desc, 0.1, Hg.desc, Sg.desc, beta, h, k, st))
    _lib.check(lib.cntrx_dyson_timestep_noalloc(n, Gb.desc, 0.1, Hg.desc, Sg.desc, beta, h, k, st))
    torch.cuda.synchronize()
    a, b = Ga.to_numpy(), Gb.to_numpy()
    for c in ("ret", "les", "tv", "mat"):
        assert np.array_equal(a[c], b[c]), (s, n, c)
    errs = component_errors(orc, Gb, after, tsteps=[n])
    assert max(errs.values()) <= 1e-12, errs


@pytest.mark.parametrize("first,second", [((1, 300, 200, 300), (1, 40, 30, 20)),
                                          ((8, 60, 40, 60), (8, 20, 16, 12)),
                                          ((4, 70, 48, 64), (2, 30, 20, 25))])
def test_noalloc_equals_dyson_timestep_twice(cntr, orc, first, second):
    from nessi_amd import _lib
    _lib.check(_lib.load().cntrx_init(0))
    _one(cntr, orc, *first, seed=sum(first))
    _one(cntr, orc, *second, seed=sum(second))
