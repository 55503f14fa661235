"""size1 outside the instantiated {1, 2, 4, 8}: 3 and 5..7 run embedded in 4 and 8, the
element-wise utilities take any size1.  Against the oracle at the caller's size (<= 1e-12
per component), the reference being generic in size1 (cntr_dyson_impl.hpp:1618-1637):
  * the whole Dyson chain (dyson_mat fixpoint and Fourier, dyson_start, dyson_timestep)
    for fermions and for bosons at mu = 0;
  * a call writes the timestep it solves and nothing else of the caller's storage;
  * batches that mix embedded and native sizes equal the single calls bit for bit;
  * force_matsubara_hermitian, bubbles with three different sizes, extrapolate_timestep
    and distance_norm2 at sizes 3, 5 and 9; solvers at size 9 return CNTRX_EUNSUPPORTED."""
import numpy as np
import pytest

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


def dyson_inputs(orc, rng, s, nt, ntau, beta, h, sig=-1, mu=0.0):
    H0, H1 = rand_herm(rng, s), rand_herm(rng, s, 0.2)
    Hb = rand_herm(rng, s)
    if sig == 1:  # bosons: every level above mu
        H0 = H0 + (0.7 - np.linalg.eigvalsh(H0).min() + mu) * np.eye(s)
        Hb = Hb + (0.4 - np.linalg.eigvalsh(Hb).min() + mu) * np.eye(s)
    H = orc.HostFunc(nt, s)
    H[-1] = H0
    for n in range(nt + 1):
        H[n] = H0 + np.sin(0.7 * n * h) * H1
    Sig = orc.HostHerm(nt, ntau, s, sig)
    orc.green_from_H_const(Sig, mu, Hb, beta, h)
    for c in ("ret", "les", "tv", "mat"):
        getattr(Sig, c)[...] *= 0.16
    return H, Sig


def assert_close(errs, tol=TOL):
    bad = {c: e for c, e in errs.items() if not e <= tol}
    assert not bad, errs


@pytest.mark.parametrize("s,sig,mu,method", [(3, -1, 0.1, 2), (3, 1, 0.0, 0), (5, -1, -0.2, 0),
                                              (6, 1, 0.0, 2), (7, -1, 0.0, 2)])
def test_dyson_chain(cntr, orc, s, sig, mu, method):
    rng = np.random.default_rng(800 + 10 * s + sig)
    nt, ntau, beta, h, k = 30, 40, 3.0, 0.04, 5
    H, Sig = dyson_inputs(orc, rng, s, nt, ntau, beta, h, sig, mu)
    X = orc.HostHerm(nt, ntau, s, sig)
    orc.dyson_mat(X, mu, H, Sig, beta, k, method)
    orc.dyson_start(X, mu, H, Sig, beta, h, k)
    for n in range(k + 1, nt + 1):
        orc.dyson_timestep(n, X, mu, H, Sig, beta, h, k)
    G = cntr.herm_matrix(nt, ntau, s, sig)
    Hg, Sg = func_to_gpu(cntr, H), to_gpu(cntr, Sig)
    cntr.dyson_mat(G, mu, Hg, Sg, beta, k, method)
    errs_mat = component_errors(orc, G, X)["mat"]
    cntr.dyson_start(G, mu, Hg, Sg, beta, h, k)
    for n in range(k + 1, nt + 1):
        cntr.dyson_timestep(n, G, mu, Hg, Sg, beta, h, k)
    errs = component_errors(orc, G, X)
    errs["mat_after_dyson_mat"] = errs_mat
    assert_close(errs)


def test_a_call_writes_its_timestep_only(cntr, orc):
    """the copy back covers exactly what the call writes: every other entry of the caller's
    storage (other timesteps, the Matsubara branch) keeps its bits"""
    rng = np.random.default_rng(900)
    s, nt, ntau, beta, h, k, n = 3, 20, 16, 2.0, 0.05, 5, 12
    H, Sig = dyson_inputs(orc, rng, s, nt, ntau, beta, h)
    X = orc.HostHerm(nt, ntau, s)
    for c in ("ret", "les", "tv", "mat"):  # arbitrary state, sentinels included
        a = getattr(X, c)
        a[...] = rng.normal(size=a.shape) + 1j * rng.normal(size=a.shape)
    G = to_gpu(cntr, X)
    cntr.dyson_timestep(n, G, 0.0, func_to_gpu(cntr, H), to_gpu(cntr, Sig), beta, h, k)
    Gh = to_host(orc, G)
    o0, o1 = n * (n + 1) // 2, (n + 1) * (n + 2) // 2
    for c in ("ret", "les"):
        a, b = getattr(Gh, c), getattr(X, c)
        assert np.array_equal(a[:o0], b[:o0]) and np.array_equal(a[o1:], b[o1:]), c
        assert not np.array_equal(a[o0:o1], b[o0:o1]), c
    t0, t1 = n * (ntau + 1), (n + 1) * (ntau + 1)
    assert np.array_equal(Gh.tv[:t0], X.tv[:t0]) and np.array_equal(Gh.tv[t1:], X.tv[t1:])
    assert np.array_equal(Gh.mat, X.mat)
    Y = X.copy()
    orc.dyson_timestep(n, Y, 0.0, H, Sig, beta, h, k)
    assert_close(component_errors(orc, G, Y, tsteps=[n]))


def test_batched_mixed_sizes_equal_single_calls(cntr, orc):
    """one batched call over sizes 3, 2, 5, 8, 3, 1, 7 (own Sigma each) equals the single
    calls bit for bit, and the oracle to 1e-12 after the whole run"""
    sizes = [3, 2, 5, 8, 3, 1, 7]
    nt, ntau, beta, h, k = 16, 20, 3.0, 0.05, 5
    rng = np.random.default_rng(901)
    Xs, Hs, Ss, Gb, Gs, Hg, Sg = [], [], [], [], [], [], []
    for i, s in enumerate(sizes):
        H, Sig = dyson_inputs(orc, rng, s, nt, ntau, beta, h)
        X = orc.HostHerm(nt, ntau, s)
        orc.dyson_mat(X, 0.0, H, Sig, beta, k, 2)
        orc.dyson_start(X, 0.0, H, Sig, beta, h, k)
        Xs.append(X)
        Hs.append(H)
        Ss.append(Sig)
        Gb.append(to_gpu(cntr, X))
        Gs.append(to_gpu(cntr, X))
        Hg.append(func_to_gpu(cntr, H))
        Sg.append(to_gpu(cntr, Sig))
    for n in range(k + 1, nt + 1):
        cntr.dyson_timestep_batched(n, Gb, 0.0, Hg, Sg, beta, h, k)
        for i in range(len(sizes)):
            cntr.dyson_timestep(n, Gs[i], 0.0, Hg[i], Sg[i], beta, h, k)
            orc.dyson_timestep(n, Xs[i], 0.0, Hs[i], Ss[i], beta, h, k)
    for i, s in enumerate(sizes):
        a, b = Gb[i].to_numpy(), Gs[i].to_numpy()
        for c in ("ret", "les", "tv", "mat"):
            assert np.array_equal(a[c], b[c]), (s, c)
        assert_close(component_errors(orc, Gb[i], Xs[i]))


def test_force_matsubara_hermitian(cntr, orc):
    rng = np.random.default_rng(902)
    X = orc.HostHerm(-1, 24, 3)
    X.mat[...] = rng.normal(size=X.mat.shape) + 1j * rng.normal(size=X.mat.shape)
    G = to_gpu(cntr, X)
    orc.force_matsubara_hermitian(X)
    cntr.force_matsubara_hermitian(G)
    assert np.abs(to_host(orc, G).mat - X.mat).max() <= 1e-15 * np.abs(X.mat).max()


@pytest.mark.parametrize("which", [1, 2])
def test_bubbles_three_sizes(cntr, orc, which):
    """C of size 2 from A of size 3 and B of size 5 (the bubbles index single elements)"""
    rng = np.random.default_rng(903 + which)
    nt, ntau, beta, h = 8, 12, 2.0, 0.1
    A = orc.HostHerm(nt, ntau, 3)
    orc.green_from_H_const(A, 0.0, rand_herm(rng, 3), beta, h)
    B = orc.HostHerm(nt, ntau, 5, 1)
    orc.green_from_H_const(B, 0.0, rand_herm(rng, 5) + 3.0 * np.eye(5), beta, h)
    X = orc.HostHerm(nt, ntau, 2)
    C = cntr.herm_matrix(nt, ntau, 2)
    Ag, Bg = to_gpu(cntr, A), to_gpu(cntr, B)
    f = orc.bubble1 if which == 1 else orc.bubble2
    g = cntr.Bubble1 if which == 1 else cntr.Bubble2
    for t in range(-1, nt + 1):
        for (c1, c2, a1, a2, b1, b2) in [(0, 1, 2, 0, 4, 1), (1, 0, 1, 2, 0, 3), (1, 1, 2, 2, 3, 4)]:
            f(t, X, c1, c2, A, A, a1, a2, B, B, b1, b2)
            g(t, C, c1, c2, Ag, Ag, a1, a2, Bg, Bg, b1, b2)
    Ch = to_host(orc, C)
    for c in ("ret", "les", "tv", "mat"):
        assert np.abs(getattr(Ch, c) - getattr(X, c)).max() <= 1e-14 * max(1.0, np.abs(getattr(X, c)).max())


def test_bubble_index_outside_the_matrix_is_an_error(cntr, orc):
    from nessi_amd import _lib
    A = cntr.herm_matrix(4, 8, 3)
    C = cntr.herm_matrix(4, 8, 2)
    with pytest.raises(RuntimeError, match="index out of range"):
        cntr.Bubble1(2, C, 0, 0, A, A, 3, 0, A, A, 0, 0)
    with pytest.raises(RuntimeError, match="index out of range"):
        cntr.Bubble2(2, C, 2, 0, A, A, 0, 0, A, A, 0, 0)
    assert _lib.load() is not None


@pytest.mark.parametrize("s", [3, 5, 9])
def test_extrapolate_and_distance_any_size(cntr, orc, s):
    rng = np.random.default_rng(904 + s)
    nt, ntau, beta, h = 16, 10, 2.0, 0.05
    A = orc.HostHerm(nt, ntau, s)
    orc.green_from_H_const(A, 0.0, rand_herm(rng, s), beta, h)
    B = orc.HostHerm(nt, ntau, s)
    orc.green_from_H_const(B, 0.1, rand_herm(rng, s), beta, h)
    Ag, Bg = to_gpu(cntr, A), to_gpu(cntr, B)
    for n in (5, 9, 15):
        orc.extrapolate_timestep(n, A, 5)
        cntr.extrapolate_timestep(n, Ag, 5)
    assert_close(component_errors(orc, Ag, A))
    for t in (-1, 0, 7, 16):
        d_gpu, d_cpu = cntr.distance_norm2(t, Ag, Bg), orc.distance_norm2(t, A, B)
        assert abs(d_gpu - d_cpu) <= 1e-12 * abs(d_cpu)


def test_solvers_above_size8_are_unsupported(cntr, orc):
    from nessi_amd import _lib
    G = cntr.herm_matrix(8, 8, 9)
    S = cntr.herm_matrix(8, 8, 9)
    H = cntr.function(8, 9)
    lib = _lib.load()
    assert lib.cntrx_dyson_timestep(7, G.desc, 0.0, H.desc, S.desc, 1.0, 0.1, 5, None) == -4
    assert lib.cntrx_dyson_mat(G.desc, 0.0, H.desc, S.desc, 1.0, 5, 2, 0, None) == -4
