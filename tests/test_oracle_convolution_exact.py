"""convolution_timestep (libcntr/cntr/cntr_convolution_impl.hpp:740-1609) at size1 2 is exact
for polynomial fields.

Gregory quadrature of order k, its start weights (poly_integration) and the Matsubara edge
corrections (rcorr) integrate polynomials of degree <= k exactly.  So if every component of
A and B is a polynomial of degree <= 2 in each argument, every integrand of the convolution
has degree <= 4 < k = 5 and the discrete convolution must equal the continuous one to
rounding, at every timestep: n < k (the start rules), n >= k, and Matsubara points near both
ends (the rcorr branches).  The continuous integrals are evaluated independently by
Gauss-Legendre quadrature of the same polynomials, with 2x2 complex coefficient matrices that
do not commute, so a block product in the wrong order, a missing conjugate transpose in a
continuation, a wrong sign of a Hermitian partner or of the -i of the mixed term fails here.

The partners Acc, Bcc hold the values that make the reference's continuations polynomial:
A^R(t,t') for t < t' is -[Acc^R(t',t)]^dagger, so Acc^R(i,j) = -P(t_j,t_i)^dagger continues
P past the diagonal; the same for the lesser components.  In these fields the integrals read
  C^R(t,t') = int_{t'}^{t} A^R(t,s) B^R(s,t') ds
  C^tv(t,tau) = int_0^t A^R(t,s) B^tv(s,tau) ds + int_0^beta A^tv(t,tau') B^M(tau'-tau) dtau'
  C^<(t,t') = int_0^t A^R(t,s) B^<(s,t') ds + int_0^t' A^<(t,s) B^A(s,t') ds
              - i int_0^beta A^tv(t,tau) B^vt(tau,t') dtau
  C^M(tau) = int_0^beta A^M(tau-tau') B^M(tau') dtau'
with B^M(x < 0) = sig B^M(beta + x), B^A(s,t) = [Bcc^R(t,s)]^dagger (= -P_BR(s,t) here) and
B^vt(tau,t) = -sig [Bcc^tv(t, beta - tau)]^dagger.
"""
import numpy as np
import pytest

NT, NTAU, BETA, H, K = 14, 14, 1.3, 0.07, 5
GL_X, GL_W = np.polynomial.legendre.leggauss(8)


class Poly2:
    """P(x, y) = sum_{a,b<=2} C[a,b] x^a y^b with s x s complex C[a,b]"""

    def __init__(self, rng, s=2):
        self.C = rng.normal(size=(3, 3, s, s)) + 1j * rng.normal(size=(3, 3, s, s))

    def __call__(self, x, y):
        xp = np.array([1.0, x, x * x])
        yp = np.array([1.0, y, y * y])
        return np.einsum("a,b,abij->ij", xp, yp, self.C)


class Poly1:
    def __init__(self, rng, s=2):
        self.C = rng.normal(size=(3, s, s)) + 1j * rng.normal(size=(3, s, s))

    def __call__(self, x):
        return self.C[0] + x * self.C[1] + x * x * self.C[2]


def integral(f, a, b):
    """int_a^b f(s) ds for a matrix-valued polynomial f of low degree (Gauss-Legendre)"""
    if b <= a:
        return 0.0
    s = 0.5 * (b - a) * GL_X + 0.5 * (b + a)
    return 0.5 * (b - a) * sum(w * f(x) for w, x in zip(GL_W, s))


def dag(m):
    return m.conj().T


def build(orc, rng, sig, s=2):
    """polynomial fields and the four operands (oracle host storage) at size s"""
    t = np.arange(NT + 1) * H
    tau = np.arange(NTAU + 1) * BETA / NTAU
    P = {n: Poly2(rng, s) for n in ("AR", "AL", "AT", "BR", "BL", "BT", "BcT")}
    P["AM"], P["BM"] = Poly1(rng, s), Poly1(rng, s)
    A, Acc, B, Bcc = (orc.HostHerm(NT, NTAU, s, sig) for _ in range(4))
    for i in range(NT + 1):
        for j in range(i + 1):
            A.ret_at(i, j)[...] = P["AR"](t[i], t[j])
            Acc.ret_at(i, j)[...] = -dag(P["AR"](t[j], t[i]))
            B.ret_at(i, j)[...] = P["BR"](t[i], t[j])
            Bcc.ret_at(i, j)[...] = -dag(P["BR"](t[j], t[i]))
            A.les_at(j, i)[...] = P["AL"](t[j], t[i])
            Acc.les_at(j, i)[...] = -dag(P["AL"](t[i], t[j]))
            B.les_at(j, i)[...] = P["BL"](t[j], t[i])
            Bcc.les_at(j, i)[...] = -dag(P["BL"](t[i], t[j]))
        for m in range(NTAU + 1):
            A.tv_at(i, m)[...] = P["AT"](t[i], tau[m])
            B.tv_at(i, m)[...] = P["BT"](t[i], tau[m])
            Bcc.tv_at(i, m)[...] = P["BcT"](t[i], tau[m])
    for m in range(NTAU + 1):
        A.mat[m] = P["AM"](tau[m])
        B.mat[m] = P["BM"](tau[m])
    return t, tau, P, A, Acc, B, Bcc


def expected(t, tau, P, sig):
    """the continuous integrals of the fields, component by component"""
    s = P["AM"].C.shape[-1]

    def BM(x):
        return P["BM"](x) if x >= 0 else sig * P["BM"](BETA + x)

    ret = np.zeros(((NT + 1) * (NT + 2) // 2, s, s), complex)
    les = np.zeros_like(ret)
    tv = np.zeros(((NT + 1) * (NTAU + 1), s, s), complex)
    for n in range(NT + 1):
        for j in range(n + 1):
            ret[n * (n + 1) // 2 + j] = integral(lambda s: P["AR"](t[n], s) @ P["BR"](s, t[j]), t[j], t[n])
            # C^<(t_j, t_n)
            v = integral(lambda s: P["AR"](t[j], s) @ P["BL"](s, t[n]), 0.0, t[j])
            v = v + integral(lambda s: P["AL"](t[j], s) @ (-P["BR"](s, t[n])), 0.0, t[n])
            bvt = lambda x: -sig * dag(P["BcT"](t[n], BETA - x))  # noqa: E731
            v = v - 1j * integral(lambda x: P["AT"](t[j], x) @ bvt(x), 0.0, BETA)
            les[n * (n + 1) // 2 + j] = v
        for m in range(NTAU + 1):
            v = integral(lambda s: P["AR"](t[n], s) @ P["BT"](s, tau[m]), 0.0, t[n])
            # the Matsubara factor is piecewise polynomial: split at tau' = tau_m
            v = v + integral(lambda x: P["AT"](t[n], x) @ BM(x - tau[m]), 0.0, tau[m])
            v = v + integral(lambda x: P["AT"](t[n], x) @ BM(x - tau[m]), tau[m], BETA)
            tv[n * (NTAU + 1) + m] = v
    mat = np.zeros((NTAU + 1, s, s), complex)
    for m in range(NTAU + 1):
        mat[m] = integral(lambda x: P["AM"](tau[m] - x) @ P["BM"](x), 0.0, tau[m]) + \
            sig * integral(lambda x: P["AM"](BETA + tau[m] - x) @ P["BM"](x), tau[m], BETA)
    return {"ret": ret, "les": les, "tv": tv, "mat": mat}


@pytest.mark.parametrize("sig", [-1, 1])
def test_convolution_size2_exact_for_polynomials(orc, sig):
    rng = np.random.default_rng(77 + sig)
    t, tau, P, A, Acc, B, Bcc = build(orc, rng, sig)
    C = orc.HostHerm(NT, NTAU, 2, sig)
    for n in range(-1, NT + 1):
        assert orc.convolution_timestep(n, C, A, Acc, B, Bcc, BETA, H, K) == 0
    X = expected(t, tau, P, sig)
    for c in ("ret", "les", "tv", "mat"):
        got, want = getattr(C, c), X[c]
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err < 1e-13, (c, err)
