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


@pytest.mark.parametrize("norb,ntau,nt,window", [(2, 24, 70, 14), (8, 32, 40, 14)])
def test_windowed_k_points_equal_whole_ones_bit_for_bit(mo, norb, ntau, nt, window):
    """The loop with herm_matrix_window k-points (the retarded rows and lesser columns of
    only the last `window` timesteps on the device, the rest streamed to pinned host memory
    for k-points 0 and 3 and dropped for the others) against the same loop on whole
    herm_matrix objects: G_loc, and every component of the streamed k-points, are the same
    bit for bit; the window slides several times.  norb 8 runs the batched steps on the
    matrix cores."""
    import torch
    nkx, beta, h, g2, k = 2, 5.0, 0.02, 0.09, 5
    runs = []
    for win in (None, window):
        L = mo.MultiOrbitalLattice(nkx, norb, beta, ntau, nt, h, g2=g2, SolveOrder=k, window=win,
                                   host_k=(0, 3))
        L.matsubara(tol=-1.0, maxiter=2)
        L.start(tol=-1.0, maxiter=2)
        for n in range(k + 1, nt + 1):
            L.step(n, tol=-1.0, maxiter=2)
        L.finish()
        torch.cuda.synchronize()
        runs.append(L)
    whole, win = runs
    assert win.G[0].r0 >= nt - window, "the window did not slide to the end"
    a, b = whole.Gloc.to_numpy(), win.Gloc.to_numpy()
    for c in ("ret", "les", "tv", "mat"):
        assert np.array_equal(a[c], b[c]), ("G_loc", c)
    for q in (0, 3):
        a, b = whole.G[q].to_numpy(), win.G[q].to_numpy()
        for c in ("ret", "les", "tv", "mat"):
            assert np.array_equal(a[c], b[c]), (q, c)
    # a k-point without a host copy keeps its whole mixed and Matsubara components and the
    # resident timesteps
    a, g = whole.G[1].to_numpy(), win.G[1]
    assert np.array_equal(a["tv"], g.tv.cpu().numpy())
    lo, hi = g.resident()
    o0, o1 = lo * (lo + 1) // 2, (hi + 1) * (hi + 2) // 2
    assert np.array_equal(a["ret"][o0:o1], g.ret_window[0:o1 - o0].cpu().numpy())
    assert np.array_equal(a["les"][o0:o1], g.les_window[0:o1 - o0].cpu().numpy())
