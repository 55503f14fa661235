"""Multi-orbital k-batch (size1 8, shared Sigma = g2 G_loc, Peierls-shifted H(k,t)) on the
GPU against the same self-consistent loop run with the oracle, <= 1e-12 per component.

Both sides run a fixed number of iterations in every loop, so that the two make the same
sequence of solves; the loop bodies follow the reference drivers
(examples/programs/gw.cpp:386-430, hubbard_chain_2b.cpp:310-322)."""
import numpy as np
import pytest

from tests.gpu_util import component_errors

pytestmark = pytest.mark.gpu
TOL = 1e-12
ITERS = {"mat": 3, "start": 2, "step": 2}


@pytest.fixture(scope="module")
def mo():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nessi_amd import _lib, multiorbital
    _lib.check(_lib.load().cntrx_init(0))
    return multiorbital


def _parts(X, t):
    return [X.mat] if t == -1 else [X.ret_row(t), X.les_col(t), X.tv_row(t)]


def _lincomb(t, C, Xs, ws):
    for i, c in enumerate(_parts(C, t)):
        c[...] = sum(w * _parts(X, t)[i] for X, w in zip(Xs, ws))


def oracle_run(orc, Hh, nk, norb, beta, ntau, nt, h, g2, k, iters=ITERS):
    """the same loop on the host, the k-points of each sweep on threads of their own (the
    oracle's calls release the GIL)"""
    from concurrent.futures import ThreadPoolExecutor
    ex = ThreadPoolExecutor(max(1, min(nk, 16)))
    G = [orc.HostHerm(nt, ntau, norb) for _ in range(nk)]
    Sig, Gloc, Gold = (orc.HostHerm(nt, ntau, norb) for _ in range(3))
    H = []
    for q in range(nk):
        f = orc.HostFunc(nt, norb)
        f.data[...] = Hh[q]
        H.append(f)
    w = [1.0 / nk] * nk

    def gloc(t):
        _lincomb(t, Gold, [Gloc], [1.0])
        _lincomb(t, Gloc, G, w)

    def each(f):
        list(ex.map(f, range(nk)))

    for _ in range(iters["mat"]):
        _lincomb(-1, Sig, [Gloc], [g2])
        each(lambda q: orc.dyson_mat(G[q], 0.0, H[q], Sig, beta, k, 2, True))
        gloc(-1)
    for _ in range(iters["start"]):
        for t in range(k + 1):
            _lincomb(t, Sig, [Gloc], [g2])
        each(lambda q: orc.dyson_start(G[q], 0.0, H[q], Sig, beta, h, k))
        for t in range(k + 1):
            gloc(t)
    for n in range(k + 1, nt + 1):
        for q in range(nk):
            orc.extrapolate_timestep(n - 1, G[q], k)
        _lincomb(n, Gloc, G, w)
        for _ in range(iters["step"]):
            _lincomb(n, Sig, [Gloc], [g2])
            each(lambda q: orc.dyson_timestep(n, G[q], 0.0, H[q], Sig, beta, h, k))
            gloc(n)
    ex.shutdown()
    return Gloc, G


def test_multiorbital_k_batch_matches_oracle(mo, orc):
    nkx, norb, beta, ntau, nt, h, g2, k = 2, 8, 20.0, 32, 14, 0.01, 0.09, 5
    L = mo.MultiOrbitalLattice(nkx, norb, beta, ntau, nt, h, g2=g2, SolveOrder=k)
    L.matsubara(tol=-1.0, maxiter=ITERS["mat"])
    L.start(tol=-1.0, maxiter=ITERS["start"])
    for n in range(k + 1, nt + 1):
        L.step(n, tol=-1.0, maxiter=ITERS["step"])
    Xloc, XG = oracle_run(orc, L.Hh, L.nk, norb, beta, ntau, nt, h, g2, k)
    errs = component_errors(orc, L.Gloc, Xloc)
    assert all(e <= TOL for e in errs.values()), ("G_loc", errs)
    for q in (0, L.nk - 1):
        errs = component_errors(orc, L.G[q], XG[q])
        assert all(e <= TOL for e in errs.values()), (q, errs)


def test_multiorbital_per_k_shape_of_the_16x16_workload(mo, orc):
    """norb 8, beta 20, ntau 600, h 0.01, Sigma = 0.09 G_loc on the 2x2 k-grid to timestep
    10 (the batched steps run on the matrix cores), against the oracle's loop"""
    iters = {"mat": 2, "start": 1, "step": 1}
    nkx, norb, beta, ntau, nt, h, g2, k = 2, 8, 20.0, 600, 10, 0.01, 0.09, 5
    L = mo.MultiOrbitalLattice(nkx, norb, beta, ntau, nt, h, g2=g2, SolveOrder=k)
    L.matsubara(tol=-1.0, maxiter=iters["mat"])
    L.start(tol=-1.0, maxiter=iters["start"])
    for n in range(k + 1, nt + 1):
        L.step(n, tol=-1.0, maxiter=iters["step"])
    Xloc, XG = oracle_run(orc, L.Hh, L.nk, norb, beta, ntau, nt, h, g2, k, iters)
    errs = component_errors(orc, L.Gloc, Xloc)
    assert all(e <= TOL for e in errs.values()), ("G_loc", errs)
    for q in (0, L.nk - 1):
        errs = component_errors(orc, L.G[q], XG[q])
        assert all(e <= TOL for e in errs.values()), (q, errs)
