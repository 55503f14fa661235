"""The embedding the library uses for size1 3 and 5..7, checked on the oracle (CPU).

The HIP kernels exist for size1 1, 2, 4 and 8; a call at size1 s = 3 or 5..7 runs at the
next size sp on arguments whose s x s blocks A become [[A, 0], [0, d 1]]: d = 0 for
contour functions, d = mu + 1 for the Hamiltonian.  The claim is that the padding rows
and columns never feed the s x s part, so that part is the size-s solution.  Here the
oracle solves the same problems at size s and at size sp on the embedded arguments; the
s x s parts must agree to rounding, the off-diagonal padding blocks must stay exactly zero
and the padding diagonal finite.  A padding value that let the padded level diverge
(d = mu for bosons: the level sits at the chemical potential) or a kernel that mixed the
two diagonal blocks would fail here.
"""
import numpy as np
import pytest

from tests import physics  # noqa: F401  (registers nothing; keeps imports uniform)

COMPS = ("ret", "les", "tv", "mat")


def _sp(s):
    return 4 if s == 3 else 8


def rand_herm(rng, s, scale=0.5):
    X = rng.normal(scale=scale, size=(s, s)) + 1j * rng.normal(scale=scale, size=(s, s))
    return 0.5 * (X + X.conj().T)


def embed_herm(orc, X, sp):
    Y = orc.HostHerm(X.nt, X.ntau, sp, X.sig)
    s = X.size1
    for c in COMPS:
        getattr(Y, c)[:, :s, :s] = getattr(X, c)
    return Y


def embed_func(orc, F, sp, d):
    Y = orc.HostFunc(F.nt, sp)
    s = F.size1
    Y.data[:, :s, :s] = F.data
    for a in range(s, sp):
        Y.data[:, a, a] = d
    return Y


def check_embedded(X, Y, comps=COMPS, tol=1e-13):
    s = X.size1
    for c in comps:
        x, y = getattr(X, c), getattr(Y, c)
        scale = max(1.0, np.abs(x).max())
        assert np.abs(y[:, :s, :s] - x).max() <= tol * scale, c
        assert not np.any(y[:, :s, s:]) and not np.any(y[:, s:, :s]), c + ": padding coupled"
        assert np.all(np.isfinite(y)), c + ": padding not finite"


def _dyson_inputs(orc, rng, s, nt, ntau, beta, h, sig, mu):
    H0 = rand_herm(rng, s)
    H1 = rand_herm(rng, s, 0.2)
    if sig == 1:  # bosons: every level above mu
        H0 = H0 + (0.4 - np.linalg.eigvalsh(H0).min() + mu + 0.3) * np.eye(s)
    H = orc.HostFunc(nt, s)
    H[-1] = H0
    for n in range(nt + 1):
        H[n] = H0 + np.sin(0.7 * n * h) * H1
    Hb = rand_herm(rng, s)
    if sig == 1:
        Hb = Hb + (0.4 - np.linalg.eigvalsh(Hb).min() + mu) * np.eye(s)
    Sig = orc.HostHerm(nt, ntau, s, sig)
    orc.green_from_H_const(Sig, mu, Hb, beta, h)
    for c in COMPS:
        getattr(Sig, c)[...] *= 0.16
    return H, Sig


@pytest.mark.parametrize("s,sig,mu,method", [(3, -1, 0.1, 2), (5, -1, -0.2, 2), (3, 1, 0.0, 2),
                                              (3, -1, 0.1, 0), (6, 1, 0.0, 0), (7, -1, 0.0, 2)])
def test_dyson_chain_embedded_equals_size_s(orc, s, sig, mu, method):
    rng = np.random.default_rng(300 + 10 * s + sig)
    nt, ntau, beta, h, k = 12, 20, 3.0, 0.05, 5
    H, Sig = _dyson_inputs(orc, rng, s, nt, ntau, beta, h, sig, mu)
    sp = _sp(s)
    Hp, Sp = embed_func(orc, H, sp, mu + 1.0), embed_herm(orc, Sig, sp)
    X = orc.HostHerm(nt, ntau, s, sig)
    Y = orc.HostHerm(nt, ntau, sp, sig)
    orc.dyson_mat(X, mu, H, Sig, beta, k, method)
    orc.dyson_mat(Y, mu, Hp, Sp, beta, k, method)
    check_embedded(X, Y, ("mat",))
    orc.dyson_start(X, mu, H, Sig, beta, h, k)
    orc.dyson_start(Y, mu, Hp, Sp, beta, h, k)
    for n in range(k + 1, nt + 1):
        orc.dyson_timestep(n, X, mu, H, Sig, beta, h, k)
        orc.dyson_timestep(n, Y, mu, Hp, Sp, beta, h, k)
    check_embedded(X, Y)


def test_dyson_timestep_from_zero_padding_history(orc):
    """The library copies the caller's size-s history in with zero padding blocks, so at
    every call the padded level starts from a zero history rather than its own solution.
    Its values then differ from a consistent padded solve, but the s x s part must not."""
    rng = np.random.default_rng(401)
    s, nt, ntau, beta, h, k, mu = 3, 14, 16, 2.0, 0.05, 4, 0.05
    H, Sig = _dyson_inputs(orc, rng, s, nt, ntau, beta, h, -1, mu)
    X = orc.HostHerm(nt, ntau, s)
    orc.dyson_mat(X, mu, H, Sig, beta, k, 2)
    orc.dyson_start(X, mu, H, Sig, beta, h, k)
    Hp, Sp = embed_func(orc, H, 4, mu + 1.0), embed_herm(orc, Sig, 4)
    for n in range(k + 1, nt + 1):
        Y = embed_herm(orc, X, 4)  # fresh copy-in at every call, as the library does
        orc.dyson_timestep(n, X, mu, H, Sig, beta, h, k)
        orc.dyson_timestep(n, Y, mu, Hp, Sp, beta, h, k)
        for comp, get in (("ret", lambda G: G.ret_row(n)), ("les", lambda G: G.les_col(n)),
                          ("tv", lambda G: G.tv_row(n))):
            x, y = get(X), get(Y)
            assert np.abs(y[:, :s, :s] - x).max() <= 1e-13 * max(1.0, np.abs(x).max()), (n, comp)
            assert not np.any(y[:, :s, s:]) and not np.any(y[:, s:, :s]), (n, comp)
            assert np.all(np.isfinite(y)), (n, comp)


@pytest.mark.parametrize("s", [3, 6])
def test_vie2_chain_embedded_equals_size_s(orc, s):
    rng = np.random.default_rng(500 + s)
    nt, ntau, beta, h, k = 12, 20, 3.0, 0.05, 5
    G0 = orc.HostHerm(nt, ntau, s)
    orc.green_from_H_const(G0, 0.0, rand_herm(rng, s), beta, h)
    F = orc.HostHerm(nt, ntau, s)
    orc.green_from_H_const(F, 0.0, rand_herm(rng, s), beta, h)
    for c in COMPS:
        getattr(F, c)[...] *= 0.3
    sp = _sp(s)
    Fp, Qp = embed_herm(orc, F, sp), embed_herm(orc, G0, sp)
    X, Y = orc.HostHerm(nt, ntau, s), orc.HostHerm(nt, ntau, sp)
    orc.vie2_mat(X, F, F, G0, beta, k, 2)
    orc.vie2_mat(Y, Fp, Fp, Qp, beta, k, 2)
    for n in range(0, nt + 1):
        orc.vie2_timestep(n, X, F, F, G0, beta, h, k, 2)
        orc.vie2_timestep(n, Y, Fp, Fp, Qp, beta, h, k, 2)
    check_embedded(X, Y)


@pytest.mark.parametrize("s", [3, 5])
def test_convolution_embedded_equals_size_s(orc, s):
    rng = np.random.default_rng(600 + s)
    nt, ntau, beta, h, k = 10, 16, 2.0, 0.05, 5
    A, B = orc.HostHerm(nt, ntau, s), orc.HostHerm(nt, ntau, s)
    orc.green_from_H_const(A, 0.0, rand_herm(rng, s), beta, h)
    orc.green_from_H_const(B, 0.0, rand_herm(rng, s), beta, h)
    sp = _sp(s)
    Ap, Bp = embed_herm(orc, A, sp), embed_herm(orc, B, sp)
    X, Y = orc.HostHerm(nt, ntau, s), orc.HostHerm(nt, ntau, sp)
    for n in range(-1, nt + 1):
        orc.convolution_timestep(n, X, A, A, B, B, beta, h, k)
        orc.convolution_timestep(n, Y, Ap, Ap, Bp, Bp, beta, h, k)
    check_embedded(X, Y)


def test_green_from_H_embedded_equals_size_s(orc):
    rng = np.random.default_rng(700)
    s, nt, ntau, beta, h, mu = 5, 10, 16, 2.0, 0.05, 0.3
    H0 = rand_herm(rng, s)
    X, Y = orc.HostHerm(nt, ntau, s), orc.HostHerm(nt, ntau, 8)
    orc.green_from_H_const(X, mu, H0, beta, h)
    Hp = np.zeros((8, 8), complex)
    Hp[:s, :s] = H0
    Hp[s:, s:] = (mu + 1.0) * np.eye(8 - s)
    orc.green_from_H_const(Y, mu, Hp, beta, h)
    check_embedded(X, Y, tol=1e-12)
