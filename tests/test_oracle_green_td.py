"""Pins of the oracle's green_from_H with a time-dependent H (commutator-free exponential
propagators CF2 and CF4:2 with polynomially interpolated H,
libcntr/cntr/cntr_equilibrium_impl.hpp:754-870, cntr_utilities_impl.hpp:285-312).

  - constant H(t): U(t) = exp(i t (mu - H0)) exactly, so both orders reproduce the
    constant-H green_from_H (an independent code path) to rounding;
  - a scalar cubic eps(t): CF4:2 over one step is the two-point Gauss-Legendre rule, exact
    for cubics, and the order-k interpolation (k >= 3) is exact too, so G^R(t,t') =
    -i exp(-i int_t'^t (eps - mu)) and G^<, G^tv follow in closed form to rounding;
    CF2 is the midpoint rule, exact for linear eps;
  - a non-commuting 2 x 2 H(t) = H0 + f(t) H1 against the same propagation at h/16: the
    error falls by 2^4 per halving of h for CF4 and by 2^2 for CF2;
  - without fixHam, H(t > k) is replaced by its order-k extrapolation: for a polynomial H of
    degree <= k that is exact, so the result equals the fixHam one to rounding, and H is
    left as the reference leaves it (overwritten by the extrapolated values);
  - the timestep form fills only timestep tstp and, like the reference, applies the
    e^{i mu h n} phases (n = 0..tstp) to U(tstp) alone: at mu = 0 it equals the whole form's
    timestep; at mu != 0 its entries are the whole form's times the leftover phase
    e^{i mu h (S - tstp + n)}, S = tstp(tstp+1)/2, checked entry by entry."""
import numpy as np
import pytest

from oracle import oracle as orc


def _tri(n):
    return n * (n + 1) // 2


def _rand_herm(rng, s, scale=1.0):
    X = rng.normal(size=(s, s)) + 1j * rng.normal(size=(s, s))
    return scale * 0.5 * (X + X.conj().T)


def _func(nt, s, vals):
    H = orc.HostFunc(nt, s)
    for t in range(-1, nt + 1):
        H[t] = vals(t)
    return H


def _max(G, X, comps=("ret", "les", "tv", "mat")):
    return max(np.abs(getattr(G, c) - getattr(X, c)).max() for c in comps)


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("s", [1, 2, 4])
def test_constant_H_equals_the_constant_form(order, s):
    rng = np.random.default_rng(10 + s + order)
    nt, ntau, beta, h, mu = 40, 30, 3.0, 0.05, 0.3
    H0 = _rand_herm(rng, s)
    A, B = orc.HostHerm(nt, ntau, s), orc.HostHerm(nt, ntau, s)
    orc.green_from_H_const(A, mu, H0, beta, h)
    orc.green_from_H_td(B, mu, _func(nt, s, lambda t: H0), beta, h, 5, order)
    assert _max(A, B) < 1e-12


def test_scalar_cubic_eps_is_exact_for_cf4():
    nt, ntau, beta, h, mu = 60, 20, 2.0, 0.03, 0.2
    c = (0.4, -0.7, 1.3, -0.9)  # eps(t) = c0 + c1 t + c2 t^2 + c3 t^3, t = n h
    eps = lambda x: c[0] + c[1] * x + c[2] * x ** 2 + c[3] * x ** 3
    ieps = lambda x: c[0] * x + c[1] * x ** 2 / 2 + c[2] * x ** 3 / 3 + c[3] * x ** 4 / 4
    for k in (3, 4, 5):
        G = orc.HostHerm(nt, ntau, 1)
        orc.green_from_H_td(G, mu, _func(nt, 1, lambda t: eps(max(t, 0) * h) if t >= 0 else eps(0.0)),
                            beta, h, k, 4)
        e0 = eps(0.0)  # H(-1): the equilibrium Hamiltonian
        w = e0 - mu
        nF = 1.0 / (1.0 + np.exp(beta * w))
        phase = lambda n: np.exp(-1j * (ieps(n * h) - mu * n * h))  # U(t) of a scalar
        err = 0.0
        for m in range(nt + 1):
            for n in range(m + 1):
                ret = -1j * phase(m) * np.conj(phase(n))
                les = 1j * phase(n) * nF * np.conj(phase(m))
                err = max(err, abs(G.ret[_tri(m) + n, 0, 0] - ret), abs(G.les[_tri(m) + n, 0, 0] - les))
            for q in range(ntau + 1):
                tau = q * beta / ntau
                tv = 1j * phase(m) * np.exp(w * tau) * nF
                err = max(err, abs(G.tv[m * (ntau + 1) + q, 0, 0] - tv))
        assert err < 1e-12, (k, err)


def test_scalar_linear_eps_is_exact_for_cf2_and_a_cubic_is_not():
    nt, ntau, beta, h, mu = 50, 10, 2.0, 0.04, 0.0
    for c3, exact in ((0.0, True), (1.5, False)):
        eps = lambda x: 0.3 - 0.8 * x + c3 * x ** 3
        ieps = lambda x: 0.3 * x - 0.4 * x ** 2 + c3 * x ** 4 / 4
        G = orc.HostHerm(nt, ntau, 1)
        orc.green_from_H_td(G, mu, _func(nt, 1, lambda t: eps(max(t, 0) * h)), beta, h, 5, 2)
        err = max(abs(G.ret[_tri(m), 0, 0] + 1j * np.exp(-1j * ieps(m * h))) for m in range(nt + 1))
        assert (err < 1e-12) == exact, (c3, err)


@pytest.mark.parametrize("order,rate", [(4, 16.0), (2, 4.0)])
def test_noncommuting_2x2_converges_at_the_order(order, rate):
    rng = np.random.default_rng(77)
    H0, H1 = _rand_herm(rng, 2), _rand_herm(rng, 2)
    T, ntau, beta, mu = 1.2, 10, 2.0, 0.1
    f = lambda x: np.sin(2.5 * x) + 0.5 * np.cos(4.0 * x)

    def run(nt):
        h = T / nt
        G = orc.HostHerm(nt, ntau, 2)
        orc.green_from_H_td(G, mu, _func(nt, 2, lambda t: H0 + f(max(t, 0) * h) * H1), beta, h, 5, order)
        return G.ret[_tri(nt):_tri(nt) + nt + 1]  # row at t = T

    ref = run(16 * 40)[::16]
    e1 = np.abs(run(40) - ref).max()
    e2 = np.abs(run(80)[::2] - ref).max()
    assert e1 > 1e-9  # the step error is visible at h = T/40
    assert 0.75 * rate < e1 / e2 < 1.3 * rate, (e1, e2, e1 / e2)


def test_extrapolated_H_equals_fixed_H_for_a_polynomial_of_degree_k():
    rng = np.random.default_rng(5)
    s, nt, ntau, beta, h, mu, k = 2, 30, 12, 2.0, 0.05, 0.0, 3
    A0, A1, A2, A3 = (_rand_herm(rng, s) for _ in range(4))
    Hp = lambda x: A0 + x * A1 + x ** 2 * A2 + x ** 3 * A3
    vals = lambda t: Hp(max(t, 0) * h)
    G1, G2 = orc.HostHerm(nt, ntau, s), orc.HostHerm(nt, ntau, s)
    orc.green_from_H_td(G1, mu, _func(nt, s, vals), beta, h, k, 4, True)
    Hw = _func(nt, s, lambda t: vals(t) if t <= k else np.zeros((s, s)))  # only H(-1..k) known
    orc.green_from_H_td(G2, mu, Hw, beta, h, k, 4, False)
    assert _max(G1, G2) < 1e-10
    assert np.abs(Hw[nt] - vals(nt)).max() < 1e-9  # H was extrapolated in place


def test_timestep_form_fills_one_timestep_with_the_reference_phase():
    rng = np.random.default_rng(9)
    s, nt, ntau, beta, h, k = 2, 20, 10, 2.0, 0.05, 5
    H0, H1 = _rand_herm(rng, s), _rand_herm(rng, s)
    vals = lambda t: H0 + np.sin(max(t, 0) * h) * H1
    for mu in (0.0, 0.4):
        W = orc.HostHerm(nt, ntau, s)
        orc.green_from_H_td(W, mu, _func(nt, s, vals), beta, h, k, 4)
        tstp = 13
        T = orc.HostHerm(nt, ntau, s)
        orc.green_from_H_td(T, mu, _func(nt, s, vals), beta, h, k, 4, True, tstp)
        # other timesteps untouched
        assert np.abs(T.ret[:_tri(tstp)]).max() == 0.0 and np.abs(T.ret[_tri(tstp + 1):]).max() == 0.0
        # U(tstp) carries the phase e^{i mu h sum_n n} instead of e^{i mu h tstp}; U(n) none
        ph_t = np.exp(1j * mu * h * (tstp * (tstp + 1) / 2 - tstp))
        for n in range(tstp + 1):
            # U(tstp) U(n)^dagger: the phases cancel on the diagonal n = tstp
            f = 1.0 if n == tstp else ph_t * np.exp(1j * mu * h * n)
            assert np.abs(T.ret[_tri(tstp) + n] - W.ret[_tri(tstp) + n] * f).max() < 1e-12
            assert np.abs(T.les[_tri(tstp) + n] - W.les[_tri(tstp) + n] * np.conj(f)).max() < 1e-12
        want_tv = W.tv[tstp * (ntau + 1):(tstp + 1) * (ntau + 1)] * ph_t
        assert np.abs(T.tv[tstp * (ntau + 1):(tstp + 1) * (ntau + 1)] - want_tv).max() < 1e-12
