"""GPU parity: every cntrx entry point against the CPU oracle on the same seeded
inputs, bar ||gpu - oracle||_max / ||oracle||_max <= 1e-12 per Keldysh component."""
import numpy as np
import pytest

from tests import physics
from tests.gpu_util import component_errors, func_to_gpu, to_gpu, to_host

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


def rand_herm(rng, s, scale=0.5):
    X = rng.normal(scale=scale, size=(s, s)) + 1j * rng.normal(scale=scale, size=(s, s))
    return 0.5 * (X + X.conj().T)


def free_gf(orc, nt, ntau, s, H0, mu, beta, h, sig=-1):
    G = orc.HostHerm(nt, ntau, s, sig)
    orc.green_from_H_const(G, mu, H0, beta, h)
    return G


def assert_close(errs, tol=TOL):
    bad = {c: e for c, e in errs.items() if not e <= tol}
    assert not bad, errs


@pytest.mark.parametrize("s", [1, 2, 3, 4, 5, 8])
def test_green_from_H(cntr, orc, s):
    rng = np.random.default_rng(10 + s)
    nt, ntau, beta, h = 12, 40, 4.0, 0.05
    H0 = rand_herm(rng, s)
    X = free_gf(orc, nt, ntau, s, H0, 0.1, beta, h)
    G = cntr.herm_matrix(nt, ntau, s)
    cntr.green_from_H(G, 0.1, H0, beta, h)
    assert_close(component_errors(orc, G, X), 1e-12)


@pytest.mark.parametrize("s", [1, 2, 3, 4, 6, 8])
def test_convolution_timestep(cntr, orc, s):
    rng = np.random.default_rng(20 + s)
    nt, ntau, beta, h, k = 14, 30, 3.0, 0.04, 5
    A = free_gf(orc, nt, ntau, s, rand_herm(rng, s), 0.0, beta, h)
    B = free_gf(orc, nt, ntau, s, rand_herm(rng, s), 0.0, beta, h)
    X = orc.HostHerm(nt, ntau, s)
    Ag, Bg = to_gpu(cntr, A), to_gpu(cntr, B)
    C = cntr.herm_matrix(nt, ntau, s)
    for n in range(-1, nt + 1):
        orc.convolution_timestep(n, X, A, A, B, B, beta, h, k)
        cntr.convolution_timestep(n, C, Ag, Ag, Bg, Bg, beta, h, k)
    assert_close(component_errors(orc, C, X))


def _rand_hh(orc, rng, nt, ntau, s, sig=-1):
    X = orc.HostHerm(nt, ntau, s, sig)
    for c in ("ret", "les", "tv", "mat"):
        a = getattr(X, c)
        a[...] = rng.normal(size=a.shape) + 1j * rng.normal(size=a.shape)
    return X


@pytest.mark.parametrize("with_f", [False, True])
@pytest.mark.parametrize("s", [1, 2, 3, 4, 7, 8])
def test_incr_convolution(cntr, orc, s, with_f):
    """C += alpha A*f*B on every timestep, random masks, with A != Acc and B != Bcc so
    that every choice between an operand and its conjugate partner is checked."""
    import torch
    rng = np.random.default_rng(40 + s + 10 * with_f)
    nt, ntau, beta, h, k = 16, 14, 3.0, 0.05, 4
    A, Acc, B, Bcc, X = (_rand_hh(orc, rng, nt, ntau, s) for _ in range(5))
    Ag, Accg, Bg, Bccg, Cg = (to_gpu(cntr, Y) for Y in (A, Acc, B, Bcc, X))
    f0 = ft = f0g = ftg = None
    if with_f:
        f0 = rng.normal(size=(s, s)) + 1j * rng.normal(size=(s, s))
        ft = rng.normal(size=(nt + 1, s, s)) + 1j * rng.normal(size=(nt + 1, s, s))
        f0g, ftg = torch.from_numpy(f0).cuda(), torch.from_numpy(ft).cuda()
    alpha = 0.7 + 0.4j
    for t in range(-1, nt + 1):
        nsel = ntau + 1 if t == -1 else 2 * (t + 1) + ntau + 1
        mask = None if t % 3 == 0 else (rng.random(nsel) < 0.6).astype(np.uint8)
        print("incr_convolution size1", s, "f", with_f, "tstp", t, flush=True)
        orc.incr_convolution(t, mask, alpha, X, A, Acc, f0, ft if t >= 0 else None, B, Bcc, beta, h, k)
        cntr.incr_convolution(t, alpha, Cg, Ag, Accg, f0g, ftg if t >= 0 else None, Bg, Bccg,
                              beta, h, k, mask=mask)
    assert_close(component_errors(orc, Cg, X))


@pytest.mark.parametrize("s", [1, 4])
def test_convolution_timestep_new_with_function(cntr, orc, s):
    """convolution_timestep_new (cntr_convolution_impl.hpp:4810-4863): C(t) = (A*f*B)(t)"""
    rng = np.random.default_rng(60 + s)
    nt, ntau, beta, h, k = 30, 24, 5.0, 0.03, 5
    A, B, C0 = (_rand_hh(orc, rng, nt, ntau, s) for _ in range(3))
    F = orc.HostFunc(nt, s)
    F.data[...] = rng.normal(size=F.data.shape) + 1j * rng.normal(size=F.data.shape)
    Ag, Bg, Cg = to_gpu(cntr, A), to_gpu(cntr, B), to_gpu(cntr, C0)
    Fg = func_to_gpu(cntr, F)
    X = orc.HostHerm(nt, ntau, s)
    for t in range(-1, nt + 1):
        fts = F.data[0] if t == -1 else F.data[1:]
        orc.incr_convolution(t, None, 1.0, X, A, A, F.data[0], fts if t >= 0 else None, B, B, beta, h, k)
        cntr.convolution_timestep_new(t, Cg, Ag, Ag, Fg, Bg, Bg, beta, h, k)
    assert_close(component_errors(orc, Cg, X))


def _c1_inputs(orc, nt=100, ntau=200, beta=5.0, h=0.02):
    # norb=1, eps(t)=0 (t<0), 0.5 sin(0.3 t) (t>=0); Sigma = V^2 g(eps_b=1), V=0.5
    H = orc.HostFunc(nt, 1)
    H[-1] = 0.0
    for n in range(nt + 1):
        H[n] = 0.5 * np.sin(0.3 * n * h)
    Sig = free_gf(orc, nt, ntau, 1, [[1.0]], 0.0, beta, h)
    for c in ("ret", "les", "tv", "mat"):
        getattr(Sig, c)[...] *= 0.25
    return H, Sig


def test_dyson_chain_config1(cntr, orc):
    """Plain single solve: dyson_mat, dyson_start, dyson_timestep through nt=100."""
    nt, ntau, beta, h, k = 100, 200, 5.0, 0.02, 5
    H, Sig = _c1_inputs(orc, nt, ntau, beta, h)
    X = orc.HostHerm(nt, ntau, 1)
    orc.dyson_mat(X, 0.0, H, Sig, beta, k, 2)
    orc.dyson_start(X, 0.0, H, Sig, beta, h, k)
    for n in range(k + 1, nt + 1):
        orc.dyson_timestep(n, X, 0.0, H, Sig, beta, h, k)
    G = cntr.herm_matrix(nt, ntau, 1)
    Hg, Sg = func_to_gpu(cntr, H), to_gpu(cntr, Sig)
    cntr.dyson_mat(G, 0.0, Hg, Sg, beta, k, cntr.CNTR_MAT_FIXPOINT)
    errs_mat = component_errors(orc, G, X)["mat"]
    cntr.dyson_start(G, 0.0, Hg, Sg, beta, h, k)
    for n in range(k + 1, nt + 1):
        cntr.dyson_timestep(n, G, 0.0, Hg, Sg, beta, h, k)
    errs = component_errors(orc, G, X)
    errs["mat_after_dyson_mat"] = errs_mat
    assert_close(errs)


@pytest.mark.parametrize("s,nt,ntau", [(1, 24, 40), (2, 24, 40), (4, 24, 40), (8, 24, 40),
                                        (2, 110, 16), (4, 110, 16), (8, 110, 16),
                                        (3, 24, 40), (5, 24, 40), (6, 24, 40), (7, 24, 40),
                                        (3, 110, 16)])
def test_dyson_single_step_from_same_state(cntr, orc, s, nt, ntau):
    """dyson_start and one dyson_timestep from identical inputs (isolates each); nt = 110
    spans several panels of the size1 >= 2 recurrences."""
    rng = np.random.default_rng(30 + s)
    beta, h, k = 4.0, 0.05, 5
    H = orc.HostFunc(nt, s)
    H0 = rand_herm(rng, s)
    H1 = rand_herm(rng, s, 0.2)
    H[-1] = H0
    for n in range(nt + 1):
        H[n] = H0 + np.sin(0.7 * n * h) * H1
    Sig = free_gf(orc, nt, ntau, s, rand_herm(rng, s), 0.0, beta, h)
    for c in ("ret", "les", "tv", "mat"):
        getattr(Sig, c)[...] *= 0.16
    X = orc.HostHerm(nt, ntau, s)
    orc.dyson_mat(X, 0.1, H, Sig, beta, k, 2)
    G = to_gpu(cntr, X)
    Hg, Sg = func_to_gpu(cntr, H), to_gpu(cntr, Sig)
    orc.dyson_start(X, 0.1, H, Sig, beta, h, k)
    cntr.dyson_start(G, 0.1, Hg, Sg, beta, h, k)
    assert_close(component_errors(orc, G, X, tsteps=range(0, k + 1)))
    for n in range(k + 1, nt + 1):
        orc.dyson_timestep(n, X, 0.1, H, Sig, beta, h, k)
    # one GPU step at every n (every 7th for the long runs) from the oracle's state
    steps = range(k + 1, nt + 1) if nt < 50 else sorted(set(range(k + 1, nt + 1, 7)) | {nt})
    for n in steps:
        G = to_gpu(cntr, X)
        cntr.dyson_timestep(n, G, 0.1, Hg, Sg, beta, h, k)
        assert_close(component_errors(orc, G, X, tsteps=[n]))


def test_dyson_size1_across_the_recurrence_variants(cntr, orc):
    """size 1 at n = 500 .. 700: the retarded recurrence keeps a 512-entry history in LDS
    below n = 512 and a 2048-entry one from n = 512 on; one GPU step at each n from the
    oracle's state, on both sides of the switch"""
    rng = np.random.default_rng(512)
    nt, ntau, beta, h, k = 700, 16, 3.0, 0.02, 5
    H = orc.HostFunc(nt, 1)
    H0, H1 = rand_herm(rng, 1), rand_herm(rng, 1, 0.2)
    H[-1] = H0
    for n in range(nt + 1):
        H[n] = H0 + np.sin(0.5 * n * h) * H1
    Sig = free_gf(orc, nt, ntau, 1, rand_herm(rng, 1), 0.0, beta, h)
    for c in ("ret", "les", "tv", "mat"):
        getattr(Sig, c)[...] *= 0.2
    X = orc.HostHerm(nt, ntau, 1)
    orc.dyson_mat(X, 0.0, H, Sig, beta, k, 2)
    orc.dyson_start(X, 0.0, H, Sig, beta, h, k)
    for n in range(k + 1, nt + 1):
        orc.dyson_timestep(n, X, 0.0, H, Sig, beta, h, k)
    Hg, Sg = func_to_gpu(cntr, H), to_gpu(cntr, Sig)
    for n in (500, 510, 511, 512, 513, 600, 700):
        G = to_gpu(cntr, X)
        cntr.dyson_timestep(n, G, 0.0, Hg, Sg, beta, h, k)
        assert_close(component_errors(orc, G, X, tsteps=[n]))
    # a batched call takes the history size of its largest n: timestep 300 of a solve batched
    # with one at n = 600 (2048-entry retarded history) against the same step alone (512
    # entries).  The two instantiations differ only in the size of that history, so ret
    # and tv agree bit for bit.  The lesser convolution differs: alone, the step sums it as the
    # one-column GEMM (k_lesconv_mm); the batch with unequal n runs k_lesconv, whose order of
    # summation differs, so les agrees to rounding (measured: 3e-16 absolute on values of
    # order 0.3)
    A, B, C = to_gpu(cntr, X), to_gpu(cntr, X), to_gpu(cntr, X)
    cntr.dyson_timestep(300, A, 0.0, Hg, Sg, beta, h, k)
    cntr.dyson_timestep_batched([300, 600], [B, C], [0.0, 0.0], [Hg, Hg], [Sg, Sg], [beta, beta],
                                [h, h], k)
    o = 300 * 301 // 2
    for a, b in ((A.ret[o:o + 301], B.ret[o:o + 301]), (A.tv[300 * 17:301 * 17], B.tv[300 * 17:301 * 17])):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    a, b = A.les[o:o + 301].cpu().numpy(), B.les[o:o + 301].cpu().numpy()
    assert np.abs(a - b).max() <= 1e-14 * np.abs(a).max()


@pytest.mark.parametrize("k", [1, 2, 3, 4])
@pytest.mark.parametrize("s,nt", [(1, 150), (2, 40)])
def test_dyson_chain_solve_orders(cntr, orc, s, nt, k):
    """dyson_mat, dyson_start and every dyson_timestep at SolveOrder 1..4 (the end-weight
    rows, BDF band and start block all change with k); nt = 150 spans two 64-row blocks
    and the partial third of the size1 = 1 recurrences."""
    rng = np.random.default_rng(70 + 10 * s + k)
    ntau, beta, h = 24, 3.0, 0.03
    H = orc.HostFunc(nt, s)
    H0, H1 = rand_herm(rng, s), rand_herm(rng, s, 0.2)
    H[-1] = H0
    for n in range(nt + 1):
        H[n] = H0 + np.sin(0.5 * n * h) * H1
    Sig = free_gf(orc, nt, ntau, s, rand_herm(rng, s), 0.0, beta, h)
    for c in ("ret", "les", "tv", "mat"):
        getattr(Sig, c)[...] *= 0.2
    X = orc.HostHerm(nt, ntau, s)
    orc.dyson_mat(X, 0.0, H, Sig, beta, k, 2)
    orc.dyson_start(X, 0.0, H, Sig, beta, h, k)
    for n in range(k + 1, nt + 1):
        orc.dyson_timestep(n, X, 0.0, H, Sig, beta, h, k)
    G = cntr.herm_matrix(nt, ntau, s)
    Hg, Sg = func_to_gpu(cntr, H), to_gpu(cntr, Sig)
    cntr.dyson_mat(G, 0.0, Hg, Sg, beta, k, cntr.CNTR_MAT_FIXPOINT)
    cntr.dyson_start(G, 0.0, Hg, Sg, beta, h, k)
    for n in range(k + 1, nt + 1):
        cntr.dyson_timestep(n, G, 0.0, Hg, Sg, beta, h, k)
    assert_close(component_errors(orc, G, X))


def test_dyson_mat_fourier(cntr, orc):
    nt, ntau, beta, h, k = -1, 200, 5.0, 0.02, 5
    H, Sig = _c1_inputs(orc, 10, ntau, beta, h)
    X = orc.HostHerm(-1, ntau, 1)
    orc.dyson_mat(X, 0.0, H, Sig, beta, k, 0)
    G = cntr.herm_matrix(-1, ntau, 1)
    Sm = cntr.herm_matrix(-1, ntau, 1)
    Sm.mat.copy_(__import__("torch").from_numpy(Sig.mat))
    cntr.dyson_mat(G, 0.0, func_to_gpu(cntr, H), Sm, beta, k, cntr.CNTR_MAT_FOURIER)
    assert_close({"mat": component_errors(orc, G, X)["mat"]})


@pytest.mark.parametrize("s", [1, 2, 3, 5, 8])
def test_vie2_chain(cntr, orc, s):
    rng = np.random.default_rng(40 + s)
    nt, ntau, beta, h, k = 30, 60, 4.0, 0.05, 5
    G0 = free_gf(orc, nt, ntau, s, rand_herm(rng, s), 0.0, beta, h)
    F = free_gf(orc, nt, ntau, s, rand_herm(rng, s), 0.0, beta, h)
    for c in ("ret", "les", "tv", "mat"):
        getattr(F, c)[...] *= 0.3
    X = orc.HostHerm(nt, ntau, s)
    orc.vie2_mat(X, F, F, G0, beta, k, 2)
    for n in range(0, nt + 1):
        orc.vie2_timestep(n, X, F, F, G0, beta, h, k, 2)
    G = cntr.herm_matrix(nt, ntau, s)
    Fg, Qg = to_gpu(cntr, F), to_gpu(cntr, G0)
    cntr.vie2_mat(G, Fg, Fg, Qg, beta, k, cntr.CNTR_MAT_FIXPOINT)
    for n in range(0, nt + 1):
        cntr.vie2_timestep(n, G, Fg, Fg, Qg, beta, h, k)
    assert_close(component_errors(orc, G, X))


@pytest.mark.parametrize("which", [1, 2])
def test_bubbles(cntr, orc, which):
    rng = np.random.default_rng(50 + which)
    nt, ntau, beta, h = 10, 20, 2.0, 0.1
    s = 2
    A = free_gf(orc, nt, ntau, s, rand_herm(rng, s), 0.0, beta, h)
    B = free_gf(orc, nt, ntau, s, rand_herm(rng, s), 0.0, beta, h, sig=1)
    X = orc.HostHerm(nt, ntau, s, -1)
    C = cntr.herm_matrix(nt, ntau, s, -1)
    Ag, Bg = to_gpu(cntr, A), to_gpu(cntr, B)
    f = orc.bubble1 if which == 1 else orc.bubble2
    g = cntr.Bubble1 if which == 1 else cntr.Bubble2
    for t in range(-1, nt + 1):
        for (c1, c2, a1, a2, b1, b2) in [(0, 0, 0, 1, 1, 0), (1, 0, 1, 1, 0, 1), (0, 1, 1, 0, 1, 1)]:
            f(t, X, c1, c2, A, A, a1, a2, B, B, b1, b2)
            g(t, C, c1, c2, Ag, Ag, a1, a2, Bg, Bg, b1, b2)
    Gh = to_host(orc, C)
    for c in ("ret", "les", "tv", "mat"):
        assert np.abs(getattr(Gh, c) - getattr(X, c)).max() <= 1e-14 * max(1.0, np.abs(getattr(X, c)).max())


def test_extrapolate_distance_lincomb(cntr, orc):
    rng = np.random.default_rng(60)
    nt, ntau, beta, h, s = 20, 16, 2.0, 0.05, 2
    A = free_gf(orc, nt, ntau, s, rand_herm(rng, s), 0.0, beta, h)
    B = free_gf(orc, nt, ntau, s, rand_herm(rng, s), 0.0, beta, h)
    Ag, Bg = to_gpu(cntr, A), to_gpu(cntr, B)
    for n in (5, 9, 15):
        orc.extrapolate_timestep(n, A, 5)
        cntr.extrapolate_timestep(n, Ag, 5)
    assert_close(component_errors(orc, Ag, A))
    for t in (-1, 0, 7, 20):
        d_gpu = cntr.distance_norm2(t, Ag, Bg)
        d_cpu = orc.distance_norm2(t, A, B)
        assert abs(d_gpu - d_cpu) <= 1e-12 * abs(d_cpu)
    C = cntr.herm_matrix(nt, ntau, s)
    for t in (-1, 3, 20):
        cntr.timestep_lincomb(t, C, [Ag, Bg], [0.25, -1.5])
    Ch = to_host(orc, C)
    for t in (3, 20):
        np.testing.assert_allclose(Ch.ret_row(t), 0.25 * A.ret_row(t) - 1.5 * B.ret_row(t), atol=1e-15)
        np.testing.assert_allclose(Ch.les_col(t), 0.25 * A.les_col(t) - 1.5 * B.les_col(t), atol=1e-15)
        np.testing.assert_allclose(Ch.tv_row(t), 0.25 * A.tv_row(t) - 1.5 * B.tv_row(t), atol=1e-15)
    np.testing.assert_allclose(Ch.mat, 0.25 * A.mat - 1.5 * B.mat, atol=1e-15)


def test_batched_equals_single(cntr, orc):
    """Batched launch over many k equals the per-k launches bit for bit."""
    nt, ntau, beta, h, k, nk = 20, 30, 4.0, 0.05, 5, 7
    Sig = free_gf(orc, nt, ntau, 1, [[0.3]], 0.0, beta, h)
    for c in ("ret", "les", "tv", "mat"):
        getattr(Sig, c)[...] *= 0.25
    Sg = to_gpu(cntr, Sig)
    Gs, Hs, Xs = [], [], []
    for q in range(nk):
        eps = -2 * np.cos(2 * np.pi * q / nk)
        H = orc.HostFunc(nt, 1)
        H.data[:] = eps
        X = orc.HostHerm(nt, ntau, 1)
        orc.dyson_mat(X, 0.0, H, Sig, beta, k, 2)
        orc.dyson_start(X, 0.0, H, Sig, beta, h, k)
        Xs.append(X)
        Gs.append(to_gpu(cntr, X))
        Hs.append(func_to_gpu(cntr, H))
    singles = [to_gpu(cntr, X) for X in Xs]
    for n in range(k + 1, nt + 1):
        cntr.dyson_timestep_batched(n, Gs, 0.0, Hs, [Sg] * nk, beta, h, k)
        for q in range(nk):
            cntr.dyson_timestep(n, singles[q], 0.0, Hs[q], Sg, beta, h, k)
    for q in range(nk):
        a, b = Gs[q].to_numpy(), singles[q].to_numpy()
        for c in ("ret", "les", "tv", "mat"):
            assert np.array_equal(a[c], b[c])


def test_no_inter_workgroup_wait_gave_up(cntr):
    """the bounded waits of the multi-workgroup recurrences never expired in this run"""
    import ctypes
    from nessi_amd import _lib
    n = ctypes.c_long(-1)
    _lib.check(_lib.load().cntrx_wait_timeouts(ctypes.byref(n)))
    assert n.value == 0


@pytest.mark.parametrize("ntau", [24, 100])
def test_dyson_batched_shared_sigma_size8_on_the_matrix_cores(cntr, orc, ntau):
    """size1 8 solves sharing Sigma (k-points of a lattice with a local self-energy) take the
    GEMM path: the Matsubara integral of the mixed component and the lesser convolution as
    complex FP64 GEMMs over the solves.  One batched GPU step at each n from the oracle's
    state (n = 6..30: start-table and Gregory weights; ntau 24: the end corrections of both
    tau intervals meet) against the oracle for every solve."""
    rng = np.random.default_rng(880 + ntau)
    s, nt, beta, h, k, nsol = 8, 30, 4.0, 0.04, 5, 3
    Sig = free_gf(orc, nt, ntau, s, rand_herm(rng, s), 0.0, beta, h)
    for c in ("ret", "les", "tv", "mat"):
        getattr(Sig, c)[...] *= 0.1
    Hs, Xs = [], []
    for q in range(nsol):
        H = orc.HostFunc(nt, s)
        H0, H1 = rand_herm(rng, s), rand_herm(rng, s, 0.2)
        H[-1] = H0
        for n in range(nt + 1):
            H[n] = H0 + np.sin((0.4 + 0.3 * q) * n * h) * H1
        X = orc.HostHerm(nt, ntau, s)
        orc.dyson_mat(X, 0.0, H, Sig, beta, k, 2)
        orc.dyson_start(X, 0.0, H, Sig, beta, h, k)
        for n in range(k + 1, nt + 1):
            orc.dyson_timestep(n, X, 0.0, H, Sig, beta, h, k)
        Hs.append(H)
        Xs.append(X)
    Sg = to_gpu(cntr, Sig)
    Hg = [func_to_gpu(cntr, H) for H in Hs]
    for n in (6, 9, 13, 20, 30):
        Gs = [to_gpu(cntr, X) for X in Xs]
        cntr.dyson_timestep_batched(n, Gs, 0.0, Hg, [Sg] * nsol, beta, h, k)
        for G, X in zip(Gs, Xs):
            assert_close(component_errors(orc, G, X, tsteps=[n]))
