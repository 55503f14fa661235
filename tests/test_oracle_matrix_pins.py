"""Pins of the oracle's matrix-valued (size1 >= 2) algorithms against closed forms.

Every system here is built from non-commuting complex matrices, so a block product in the
wrong order, a missing conjugation or a transposed orbital index changes the answer at
O(1) of the couplings, far above the tolerances asserted:
  * green_from_H for 2x2 and 4x4 Hamiltonians against the eigen-decomposition formulas;
  * Bubble1/Bubble2 at size 2 for every orbital index map, against the two-time products
    of closed-form free GFs (the reference's test/bubble.cpp:818-1161 does the same with
    one index map and real rotations);
  * convolution_timestep at size 2: A (Hb - Ha) * B = B - A for free GFs of Ha, Hb
    (the matrix form of the single-pole identity of test/convolution.cpp:20-27);
  * dyson_mat/dyson_start/dyson_timestep and vie2_* at size 2: a 2-level system coupled to
    a 2-level bath, G_ss of the exact 4x4 free GF (the matrix form of test/dyson.cpp:23-128),
    plus the order of convergence in h of the retarded component.
"""
import numpy as np
import pytest

from tests import physics

HA = np.array([[0.7, 0.3 + 0.2j], [0.3 - 0.2j, -0.4]])
HB = np.array([[-0.5, -0.25j], [0.25j, 0.9]])


def _herm(orc, nt, ntau, s, sig=-1):
    return orc.HostHerm(nt, ntau, s, sig)


def _comp_err(G, X, comps=("ret", "les", "tv", "mat")):
    return {c: float(np.abs(getattr(G, c) - getattr(X, c)).max()) for c in comps}


@pytest.mark.parametrize("s,sig", [(2, -1), (4, -1), (2, 1)])
def test_green_from_H_matrix_closed_form(orc, s, sig):
    rng = np.random.default_rng(11 + s)
    X = rng.normal(size=(s, s)) + 1j * rng.normal(size=(s, s))
    Hm = 0.3 * (X + X.conj().T)
    mu = -0.2
    if sig == 1:  # bosons: every level above mu
        Hm = Hm + (0.5 - np.linalg.eigvalsh(Hm).min() + mu) * np.eye(s)
    nt, ntau, beta, h = 25, 60, 3.0, 0.05
    G = _herm(orc, nt, ntau, s, sig)
    assert orc.green_from_H_const(G, mu, Hm, beta, h) == 0
    Y = physics.free_matrix(_herm(orc, nt, ntau, s, sig), Hm, mu, beta, h, sig)
    err = _comp_err(G, Y)
    assert max(err.values()) < 1e-13, err


def _index_map(c1, c2):
    # every C element reads different A and B elements, none of them diagonal-only
    return (c2, 1 - c1), (1 - c2, c1)


@pytest.mark.parametrize("which", [1, 2])
@pytest.mark.parametrize("sa,sb", [(-1, -1), (-1, 1), (1, -1), (1, 1)])
def test_bubble_size2_every_index_map(orc, which, sa, sb):
    nt, ntau, beta, h = 10, 16, 1.5, 0.1

    def ham(sig, Hm):
        return Hm + (0.6 - np.linalg.eigvalsh(Hm).min()) * np.eye(2) if sig == 1 else Hm

    Ha, Hb = ham(sa, HA), ham(sb, HB)
    A = physics.free_matrix(_herm(orc, nt, ntau, 2, sa), Ha, 0.0, beta, h, sa)
    B = physics.free_matrix(_herm(orc, nt, ntau, 2, sb), Hb, 0.0, beta, h, sb)
    fa = physics.free_matrix_full(Ha, 0.0, beta, h, nt, ntau, sa)
    fb = physics.free_matrix_full(Hb, 0.0, beta, h, nt, ntau, sb)
    C = _herm(orc, nt, ntau, 2, sa * sb)
    f = orc.bubble1 if which == 1 else orc.bubble2
    for t in range(-1, nt + 1):
        for c1 in range(2):
            for c2 in range(2):
                (a1, a2), (b1, b2) = _index_map(c1, c2)
                assert f(t, C, c1, c2, A, A, a1, a2, B, B, b1, b2) == 0
    worst = 0.0
    for c1 in range(2):
        for c2 in range(2):
            (a1, a2), (b1, b2) = _index_map(c1, c2)
            Al, Ag = fa["les"][:, :, a1, a2], fa["gtr"][:, :, a1, a2]
            if which == 1:
                # C(t,t') = i A_{a1a2}(t,t') B_{b2b1}(t',t)
                Bl, Bg = fb["les"][:, :, b2, b1].T, fb["gtr"][:, :, b2, b1].T
                ctv = 1j * fa["tv"][:, :, a1, a2] * fb["vt"][:, :, b2, b1].T
                cm = -fa["mat"][:, a1, a2] * fb["matm"][:, b2, b1]
                cl, cg = 1j * Al * Bg, 1j * Ag * Bl
            else:
                # C(t,t') = i A_{a1a2}(t,t') B_{b1b2}(t,t')
                Bl, Bg = fb["les"][:, :, b1, b2], fb["gtr"][:, :, b1, b2]
                ctv = 1j * fa["tv"][:, :, a1, a2] * fb["tv"][:, :, b1, b2]
                cm = -fa["mat"][:, a1, a2] * fb["mat"][:, b1, b2]
                cl, cg = 1j * Al * Bl, 1j * Ag * Bg
            for t in range(nt + 1):
                for tp in range(t + 1):
                    worst = max(worst, abs(C.ret_at(t, tp)[c1, c2] - (cg[t, tp] - cl[t, tp])),
                                abs(C.les_at(tp, t)[c1, c2] - cl[tp, t]))
                worst = max(worst, np.abs(C.tv_row(t)[:, c1, c2] - ctv[t]).max())
            worst = max(worst, np.abs(C.mat[:, c1, c2] - cm).max())
    assert worst < 1e-13, worst


def _right(X, D):
    """herm_matrix X with every block multiplied by the constant matrix D on the right"""
    Y = X.copy()
    for c in ("ret", "les", "tv", "mat"):
        getattr(Y, c)[...] = getattr(X, c) @ D
    return Y


def _left(D, X):
    Y = X.copy()
    for c in ("ret", "les", "tv", "mat"):
        getattr(Y, c)[...] = D @ getattr(X, c)
    return Y


@pytest.mark.parametrize("sig", [-1, 1])
def test_convolution_size2_matrix_identity(orc, sig):
    # (i d/dt - Ha) A = 1, (i d/dt - Hb) B = 1  =>  A (Hb - Ha) * B = B - A on the contour.
    # The partner of A' = A D is A'^cc = D A (A is its own partner, D Hermitian).
    nt, ntau, beta, h = 30, 500, 5.0, 0.01
    Ha, Hb = HA, HB
    if sig == 1:
        Ha = Ha + (0.5 - np.linalg.eigvalsh(Ha).min()) * np.eye(2)
        Hb = Hb + (0.5 - np.linalg.eigvalsh(Hb).min()) * np.eye(2)
    A = physics.free_matrix(_herm(orc, nt, ntau, 2, sig), Ha, 0.0, beta, h, sig)
    B = physics.free_matrix(_herm(orc, nt, ntau, 2, sig), Hb, 0.0, beta, h, sig)
    D = Hb - Ha
    assert np.abs(Ha @ Hb - Hb @ Ha).max() > 0.1  # the test needs non-commuting matrices
    Ap, Apcc = _right(A, D), _left(D, A)
    C = _herm(orc, nt, ntau, 2, sig)
    for n in range(-1, nt + 1):
        assert orc.convolution_timestep(n, C, Ap, Apcc, B, B, beta, h, 5) == 0
    X = _herm(orc, nt, ntau, 2, sig)
    for c in ("ret", "les", "tv", "mat"):
        getattr(X, c)[...] = getattr(B, c) - getattr(A, c)
    err = _comp_err(C, X)
    # the exponentials are integrated to rounding at this grid (about 1e-14)
    assert max(err.values()) < 1e-13, err
    # the products in the other order fail by far
    Cw = _herm(orc, nt, ntau, 2, sig)
    for n in range(-1, nt + 1):
        orc.convolution_timestep(n, Cw, Apcc, Ap, B, B, beta, h, 5)
    assert max(_comp_err(Cw, X).values()) > 1e-2


HS = np.array([[-1.0, 0.3 + 0.2j], [0.3 - 0.2j, 0.5]])
HBATH = np.array([[0.8, -0.25j], [0.25j, -0.4]])
VC = np.array([[0.3, 0.1 + 0.2j], [-0.15j, 0.25]])


def _embedded(orc, nt, ntau, h, beta, mu=-0.1):
    """(G_ss exact, Sigma = V g_bath V^+, H_s as a function) for the 2+2 level system"""
    H4 = np.block([[HS, VC], [VC.conj().T, HBATH]])
    G4 = physics.free_matrix(_herm(orc, nt, ntau, 4), H4, mu, beta, h)
    Gx = _herm(orc, nt, ntau, 2)
    for c in ("ret", "les", "tv", "mat"):
        getattr(Gx, c)[...] = getattr(G4, c)[:, :2, :2]
    gb = physics.free_matrix(_herm(orc, nt, ntau, 2), HBATH, mu, beta, h)
    Sig = _herm(orc, nt, ntau, 2)
    for c in ("ret", "les", "tv", "mat"):
        getattr(Sig, c)[...] = VC @ getattr(gb, c) @ VC.conj().T
    Hf = orc.HostFunc(nt, 2)
    Hf.data[...] = HS
    return Gx, Sig, Hf, mu


def _dyson_solve(orc, Sig, Hf, mu, beta, h, nt, ntau, k=5):
    G = _herm(orc, nt, ntau, 2)
    assert orc.dyson_mat(G, mu, Hf, Sig, beta, k, 2) == 0
    orc.dyson_start(G, mu, Hf, Sig, beta, h, k)
    for n in range(k + 1, nt + 1):
        orc.dyson_timestep(n, G, mu, Hf, Sig, beta, h, k)
    return G


def test_dyson_size2_embedding_exact(orc):
    nt, ntau, h, beta = 100, 400, 0.05, 5.0
    Gx, Sig, Hf, mu = _embedded(orc, nt, ntau, h, beta)
    G = _dyson_solve(orc, Sig, Hf, mu, beta, h, nt, ntau)
    err = _comp_err(G, Gx)
    # measured: about 2.5e-8 on the time components, 2e-13 on the Matsubara one
    assert max(err[c] for c in ("ret", "les", "tv")) < 1e-7, err
    assert err["mat"] < 1e-11, err


def test_dyson_size2_retarded_order_in_h(orc):
    # G^R only involves Sigma^R and H: its error at the final time T = 3 must fall like
    # h^6 (within a margin) when h is halved
    errs = []
    for nt, h in ((30, 0.1), (60, 0.05)):
        Gx, Sig, Hf, mu = _embedded(orc, nt, 50, h, 5.0)
        G = _dyson_solve(orc, Sig, Hf, mu, 5.0, h, nt, 50)
        errs.append(np.abs(G.ret_row(nt) - Gx.ret_row(nt)).max())
    # measured: 8.7e-7 -> 1.35e-8, a factor 64 (order 6)
    assert errs[1] < errs[0] / 2 ** 5.5, errs
    assert errs[1] < 3e-8, errs


def test_vie2_size2_embedding_exact(orc):
    # [1 - G0*Sigma] G = G0 with G0 the free GF of H_s: F = -G0*Sigma, F^cc = -Sigma*G0
    nt, ntau, h, beta = 60, 400, 0.05, 5.0
    Gx, Sig, Hf, mu = _embedded(orc, nt, ntau, h, beta)
    G0 = physics.free_matrix(_herm(orc, nt, ntau, 2), HS, mu, beta, h)
    F, Fcc = _herm(orc, nt, ntau, 2), _herm(orc, nt, ntau, 2)

    def conv(t):
        orc.convolution_timestep(t, F, G0, G0, Sig, Sig, beta, h, 5)
        orc.convolution_timestep(t, Fcc, Sig, Sig, G0, G0, beta, h, 5)
        for X in (F, Fcc):
            if t == -1:
                X.mat[...] *= -1
            else:
                X.ret_row(t)[...] *= -1
                X.les_col(t)[...] *= -1
                X.tv_row(t)[...] *= -1

    G = _herm(orc, nt, ntau, 2)
    conv(-1)
    assert orc.vie2_mat(G, F, Fcc, G0, beta, 5, 2) == 0
    for t in range(0, 6):
        conv(t)
    orc.vie2_start(G, F, Fcc, G0, beta, h, 5)
    for t in range(6, nt + 1):
        conv(t)
        orc.vie2_timestep(t, G, F, Fcc, G0, beta, h, 5)
    err = _comp_err(G, Gx)
    # measured: about 1.3e-10 on the time components, 2e-13 on the Matsubara one
    assert max(err[c] for c in ("ret", "les", "tv")) < 1e-9, err
    assert err["mat"] < 1e-11, err
