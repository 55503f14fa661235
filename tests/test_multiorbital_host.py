"""The multi-orbital k-batch's Hamiltonians (host side, no GPU)."""
import numpy as np

from nessi_amd import multiorbital as mo


def test_hamiltonians_follow_the_workload_definition():
    """H(k,t) = H_rand + eps(k - A(t)) 1 with A = 0 on the Matsubara branch"""
    nkx, norb, nt, h = 4, 8, 20, 0.05
    Hh = mo.hamiltonians(nkx, norb, nt, h, seed=7)
    Hr = mo.random_hermitian(norb, 7)
    assert np.allclose(Hr, Hr.conj().T)
    q, kx, ky = 5, 2 * np.pi * 1 / nkx, 2 * np.pi * 1 / nkx
    for t in (-1, 0, 7, nt):
        A = 0.0 if t == -1 else 0.5 * np.sin(2.0 * t * h) * np.exp(-(t * h - 5.0) ** 2 / 2)
        eps = -2 * (np.cos(kx - A) + np.cos(ky - A))
        assert np.allclose(Hh[q, t + 1], Hr + eps * np.eye(norb))
