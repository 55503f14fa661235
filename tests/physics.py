"""Closed-form Green's functions used to pin the oracle (no reference code involved).

Free single-level fermion/boson GF on the Kadanoff-Baym contour (units hbar=1):
  G^R(t,t')  = -i e^{-i e (t-t')}
  G^<(t,t')  = -sig*i n(e) e^{-i e (t-t')}            (n = Fermi or Bose)
  G^tv(t,tau)= -sig*i ... see free_single_pole()
  G^M(tau)   = -e^{-e tau} (1 + sig n(e))
with e = eps - mu.  These are the textbook formulas; the oracle's green_from_H must
reproduce them (libcntr/cntr/cntr_equilibrium_impl.hpp:873-953 builds the same
objects from an eigen-decomposition).
"""
import numpy as np


def fermi(beta, w):
    return 0.5 * (1.0 - np.tanh(0.5 * beta * w))


def bose(beta, w):
    return 1.0 / np.expm1(beta * w)


def free_single_pole(H, e, beta, h, sig=-1):
    """Fill a 1x1 HostHerm H with the free GF of level energy e (already minus mu)."""
    nt, ntau = H.nt, H.ntau
    n = fermi(beta, e) if sig == -1 else bose(beta, e)
    taus = np.arange(ntau + 1) * beta / ntau
    H.mat[:, 0, 0] = -np.exp(-e * taus) * (1.0 + sig * n)
    if nt < 0:
        return H
    t = np.arange(nt + 1) * h
    for i in range(nt + 1):
        j = np.arange(i + 1)
        H.ret_row(i)[:, 0, 0] = -1j * np.exp(-1j * e * (t[i] - t[j]))
        H.les_col(i)[:, 0, 0] = -sig * 1j * n * np.exp(-1j * e * (t[j] - t[i]))
        H.tv_row(i)[:, 0, 0] = -sig * 1j * n * np.exp(-1j * e * t[i]) * np.exp(e * taus)
    return H


def full_two_time(H):
    """Return (les, gtr) arrays [t,t'] of a scalar HostHerm using the Hermitian
    symmetries les(t,t') = -conj(les(t',t)), ret = gtr - les for t >= t'."""
    nt = H.nt
    les = np.zeros((nt + 1, nt + 1), complex)
    ret = np.zeros((nt + 1, nt + 1), complex)
    for i in range(nt + 1):
        for j in range(i + 1):
            les[j, i] = H.les_at(j, i)[0, 0]
            les[i, j] = -np.conj(H.les_at(j, i)[0, 0])
            ret[i, j] = H.ret_at(i, j)[0, 0]
    gtr = np.zeros_like(les)
    for i in range(nt + 1):
        for j in range(nt + 1):
            if i >= j:
                gtr[i, j] = ret[i, j] + les[i, j]
            else:
                gtr[i, j] = -np.conj(ret[j, i] + les[j, i])
    return les, gtr


def bethe_mat(beta, ntau, nquad=400):
    """G^M(tau) for the semicircular DOS of half-bandwidth 2 (mu=0):
    G(tau) = -int dw rho(w) e^{-w tau} (1 - f(w)),  rho(w) = sqrt(4-w^2)/(2 pi).
    Gauss-Legendre in w = 2 cos(theta) makes the integrand smooth."""
    x, wq = np.polynomial.legendre.leggauss(nquad)
    th = 0.5 * np.pi * (x + 1.0)
    wth = 0.5 * np.pi * wq
    w = 2.0 * np.cos(th)
    rho_jac = (2.0 * np.sin(th)) ** 2 / (2.0 * np.pi)  # rho(w) |dw/dtheta|
    taus = np.arange(ntau + 1) * beta / ntau
    g = np.empty(ntau + 1)
    for m, tau in enumerate(taus):
        # e^{-w tau}(1-f(w)) = e^{-w tau} / (1 + e^{-beta w}), evaluated stably
        val = np.where(w > 0, np.exp(-w * tau) / (1.0 + np.exp(-beta * w)),
                       np.exp((beta - tau) * w) / (1.0 + np.exp(beta * w)))
        g[m] = -np.sum(wth * rho_jac * val)
    return g


def _eig(Hm, mu):
    e, Um = np.linalg.eigh(np.asarray(Hm, np.complex128) - mu * np.eye(len(Hm)))
    return e, Um


def _occ(beta, e, sig):
    return fermi(beta, e) if sig == -1 else bose(beta, e)


def free_matrix(G, Hm, mu, beta, h, sig=-1):
    """Fill an s x s HostHerm G with the free GF of the constant Hermitian matrix Hm, from
    its eigen-decomposition H - mu = U diag(e) U^dagger (textbook formulas, every
    component evaluated directly, no contour symmetry used):
      G^R(t,t')  = -i U e^{-ie(t-t')} U^+
      G^<(t,t')  = -sig i U n(e) e^{-ie(t-t')} U^+
      G^tv(t,tau)= -sig i U n(e) e^{-iet} e^{e tau} U^+
      G^M(tau)   = -U e^{-e tau} (1 + sig n(e)) U^+"""
    e, Um = _eig(Hm, mu)
    n = _occ(beta, e, sig)
    Ud = Um.conj().T
    nt, ntau = G.nt, G.ntau
    taus = np.arange(ntau + 1) * beta / ntau

    def rot(d):  # U diag(d) U^+ for a stack of diagonals d[..., s]
        return np.einsum("ab,...b,bc->...ac", Um, d, Ud)

    G.mat[...] = rot(-np.exp(-np.outer(taus, e)) * (1.0 + sig * n)[None, :])
    if nt < 0:
        return G
    t = np.arange(nt + 1) * h
    for i in range(nt + 1):
        j = np.arange(i + 1)
        ph = np.exp(-1j * np.outer(t[i] - t[j], e))
        G.ret_row(i)[...] = rot(-1j * ph)
        phl = np.exp(-1j * np.outer(t[j] - t[i], e))
        G.les_col(i)[...] = rot(-sig * 1j * n[None, :] * phl)
        G.tv_row(i)[...] = rot(-sig * 1j * n[None, :] * np.exp(-1j * e * t[i])[None, :] *
                               np.exp(np.outer(taus, e)))
    return G


def free_matrix_full(Hm, mu, beta, h, nt, ntau, sig=-1):
    """Every component of the free GF of Hm on the whole grid, evaluated directly:
    les[t, t'], gtr[t, t'] (t, t' = 0..nt), tv[t, m], vt[m, t] (the mixed component
    G^(tau_m, t)) and mat[m] = G^M(tau_m), matm[m] = G^M(-tau_m)."""
    e, Um = _eig(Hm, mu)
    n = _occ(beta, e, sig)
    Ud = Um.conj().T
    taus = np.arange(ntau + 1) * beta / ntau
    t = np.arange(nt + 1) * h

    def rot(d):
        return np.einsum("ab,...b,bc->...ac", Um, d, Ud)

    dt = t[:, None, None] - t[None, :, None]
    ph = np.exp(-1j * dt * e[None, None, :])
    out = {
        "les": rot(-sig * 1j * n * ph),
        "gtr": rot(-1j * (1.0 + sig * n) * ph),
        "tv": rot(-sig * 1j * n * np.exp(-1j * t[:, None, None] * e) * np.exp(taus[None, :, None] * e)),
        "vt": rot(-1j * (1.0 + sig * n) * np.exp(-taus[:, None, None] * e) *
                  np.exp(1j * t[None, :, None] * e)),
        "mat": rot(-np.exp(-np.outer(taus, e)) * (1.0 + sig * n)),
        "matm": rot(-sig * n * np.exp(np.outer(taus, e))),
    }
    return out
