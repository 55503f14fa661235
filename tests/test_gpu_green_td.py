"""green_from_H with a time-dependent H (libcntr/cntr/cntr_equilibrium_impl.hpp:754-870,
955-1040) through the C ABI against the oracle, <= 1e-12 per component: size1 1, 2, 8 and
the embedded 3 and 5, both commutator-free orders, H known for all times (fixHam) or
extrapolated in place, and the timestep form at mu != 0 (the reference's phase convention)."""
import numpy as np
import pytest

from tests.gpu_util import component_errors, func_to_gpu, to_host

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def cntr():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nessi_amd import cntr as c
    from nessi_amd import _lib
    _lib.check(_lib.load().cntrx_init(0))
    return c


def _rand_herm(rng, s, scale=0.6):
    X = rng.normal(scale=scale, size=(s, s)) + 1j * rng.normal(scale=scale, size=(s, s))
    return 0.5 * (X + X.conj().T)


def _H(orc, rng, nt, s):
    H0, H1 = _rand_herm(rng, s), _rand_herm(rng, s, 0.3)
    F = orc.HostFunc(nt, s)
    F[-1] = H0
    for t in range(nt + 1):
        F[t] = H0 + np.sin(1.3 * t * 0.04) * H1 + 0.2 * np.cos(0.5 * t * 0.04) * H1 @ H1
    return F


@pytest.mark.parametrize("s", [1, 2, 3, 5, 8])
@pytest.mark.parametrize("cf", [2, 4])
def test_green_from_H_time_dependent_matches_oracle(cntr, orc, s, cf):
    rng = np.random.default_rng(300 + 10 * s + cf)
    nt, ntau, beta, h, mu, k = 60, 40, 4.0, 0.04, 0.25, 5
    F = _H(orc, rng, nt, s)
    X = orc.HostHerm(nt, ntau, s)
    orc.green_from_H_td(X, mu, F, beta, h, k, cf, True)
    G = cntr.herm_matrix(nt, ntau, s)
    cntr.green_from_H(G, mu, func_to_gpu(cntr, F), beta, h, k, cf)
    errs = component_errors(orc, G, X)
    assert all(e <= TOL for e in errs.values()), errs


@pytest.mark.parametrize("k", [2, 5])
def test_green_from_H_extrapolated_H_matches_oracle(cntr, orc, k):
    """fixHam = false: H(t > k) is replaced by its extrapolation, on both sides, and the
    caller's function holds the extrapolated values afterwards.  Each extrapolated H(t) feeds
    the next extrapolation, a recursion with the k+1-fold root 1 that amplifies rounding like
    t^k, so the propagation is kept to nt = k + 9 (the reference's use: the predictor of one
    timestep, corrected afterwards)."""
    rng = np.random.default_rng(41 + k)
    s, ntau, beta, h, mu = 2, 30, 3.0, 0.05, 0.1
    nt = k + 9
    F = _H(orc, rng, nt, s)
    Fo = orc.HostFunc(nt, s)
    Fo.data[...] = F.data
    X = orc.HostHerm(nt, ntau, s)
    orc.green_from_H_td(X, mu, Fo, beta, h, k, 4, False)
    G = cntr.herm_matrix(nt, ntau, s)
    f = func_to_gpu(cntr, F)
    cntr.green_from_H(nt, G, mu, f, beta, h, False, k, 4)  # the timestep form at nt, fixHam false
    X2 = orc.HostHerm(nt, ntau, s)
    F2 = orc.HostFunc(nt, s)
    F2.data[...] = F.data
    orc.green_from_H_td(X2, mu, F2, beta, h, k, 4, False, nt)
    errs = component_errors(orc, G, X2, tsteps=[nt])
    assert all(e <= TOL for e in errs.values()), errs
    assert np.abs(f.data.cpu().numpy() - F2.data).max() < 1e-13  # H extrapolated in place
    assert np.abs(F2.data[nt + 1] - F.data[nt + 1]).max() > 1e-12  # ... and it did change


@pytest.mark.parametrize("s", [8, 5])
def test_green_from_H_timestep_form_matches_oracle(cntr, orc, s):
    rng = np.random.default_rng(5)
    nt, ntau, beta, h, mu, k = 30, 24, 4.0, 0.03, 0.35, 5
    F = _H(orc, rng, nt, s)
    f = func_to_gpu(cntr, F)
    for tstp in (-1, 0, 3, 17, nt):
        X = orc.HostHerm(nt, ntau, s)
        orc.green_from_H_td(X, mu, F, beta, h, k, 4, True, tstp)
        G = cntr.herm_matrix(nt, ntau, s)
        cntr.green_from_H(tstp, G, mu, f, beta, h, True, k, 4)
        Gh = to_host(orc, G)
        if tstp == -1:
            assert np.abs(Gh.mat - X.mat).max() <= TOL * np.abs(X.mat).max()
            assert not np.any(Gh.ret) and not np.any(Gh.tv)
            continue
        errs = component_errors(orc, G, X, tsteps=[tstp])
        assert all(e <= TOL for e in errs.values()), (tstp, errs)
        if tstp >= 0:  # nothing else written
            o = tstp * (tstp + 1) // 2
            assert not np.any(Gh.ret[:o]) and not np.any(Gh.mat)
