"""Synthetic, seeded request streams for the batch server.

`uneven_batch` draws independent single-site solves: size1 uniform in {1,2,4,8}, nt
uniform in [nt_lo, nt_hi], beta uniform in [2,20], H(t) = H0 + 0.2 sin(Omega t) H1 with
seeded random Hermitian H0, H1 (entries N(0, 0.5^2), then (X + X^dagger)/2) and
Omega uniform in [0.5, 3]; H(-1) = H0.  The bath has one level per orbital,
eps_b uniform in [-2, 2], coupled with V = 0.4 (Sigma = V^2 g[diag(eps_b)]).
h = 0.02, ntau = 300, SolveOrder 5, mu = 0, fermions.
"""
import numpy as np

from ._lib import CNTRX_MAT_FIXPOINT


def make_request(size1, nt, ntau, beta, h, H, Hbath, V, mu=0.0, sig=-1, SolveOrder=5,
                 mat_method=CNTRX_MAT_FIXPOINT):
    """request dict; H: complex (nt+2, size1, size1) = H(t), t = -1..nt; Hbath: (size1, size1)"""
    H = np.ascontiguousarray(H, np.complex128).reshape(nt + 2, size1, size1)
    Hbath = np.ascontiguousarray(Hbath, np.complex128).reshape(size1, size1)
    return {"size1": size1, "nt": nt, "ntau": ntau, "beta": beta, "h": h, "H": H, "Hbath": Hbath,
            "V": V, "mu": mu, "sig": sig, "SolveOrder": SolveOrder, "mat_method": mat_method}


def rand_herm(rng, s, scale=0.5):
    X = rng.normal(scale=scale, size=(s, s)) + 1j * rng.normal(scale=scale, size=(s, s))
    return 0.5 * (X + X.conj().T)


def driven_H(H0, H1, amp, omega, nt, h):
    """H(t) for t = -1..nt: H(-1) = H0, H(n) = H0 + amp sin(omega n h) H1"""
    s = H0.shape[0]
    H = np.empty((nt + 2, s, s), np.complex128)
    H[0] = H0
    tt = np.arange(nt + 1) * h
    H[1:] = H0[None] + amp * np.sin(omega * tt)[:, None, None] * H1[None]
    return H


def uneven_batch(count=64, seed=2024, nt_lo=200, nt_hi=1500, ntau=300, h=0.02, sizes=(1, 2, 4, 8),
                 SolveOrder=5):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        s = int(rng.choice(sizes))
        nt = int(rng.integers(nt_lo, nt_hi + 1))
        beta = float(rng.uniform(2.0, 20.0))
        omega = float(rng.uniform(0.5, 3.0))
        H0, H1 = rand_herm(rng, s), rand_herm(rng, s)
        eb = rng.uniform(-2.0, 2.0, size=s)
        out.append(make_request(s, nt, ntau, beta, h, driven_H(H0, H1, 0.2, omega, nt, h),
                                np.diag(eb).astype(np.complex128), 0.4, SolveOrder=SolveOrder))
    return out


def poisson_arrivals(rate, count, seed=7):
    """arrival times (seconds from 0) of a Poisson process of the given rate"""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=count))
