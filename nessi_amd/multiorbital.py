"""Multi-orbital k-batch with a local self-energy fed back from G_loc, on one GPU.

The workload: a 2D square lattice on an nkx x nkx k-grid, norb orbitals per k-point,
    H(k, t) = H_rand + [-2 (cos(kx - A(t)) + cos(ky - A(t)))] 1,
    A(t) = E0 sin(omega t) exp(-(t - 5)^2 / 2)   (A = 0 on the Matsubara branch),
with H_rand a seeded random Hermitian norb x norb matrix (entries N(0, 0.5^2), then
(X + X^dagger)/2), and Sigma = g2 G_loc shared by every k (local hybridization feedback).

The driver is the reference's self-consistent k-loop (examples/programs/gw.cpp:386-430,
hubbard_chain_2b.cpp:310-322) with one GPU holding every k-point:
  Matsubara: Sigma^M = g2 G_loc^M, dyson_mat per k, G_loc^M = mean over k, until the
             distance of successive G_loc^M is below tol;
  start:     Sigma(0..K) = g2 G_loc(0..K), dyson_start per k, G_loc(0..K), until converged;
  timestep:  extrapolate_timestep per k, G_loc(n); then Sigma(n) = g2 G_loc(n), every k's
             dyson_timestep in one batched launch sequence (shared Sigma), G_loc(n), until
             the distance is below tol.
Each call is a cntrx_* entry point (cntr_dyson_decl.hpp, cntr_utilities_decl.hpp).

With window=W the k-points' Green functions are herm_matrix_window objects: the whole
G^tv history per k-point stays on the device, the retarded rows and lesser columns only
for the last W timesteps; rows leaving the window go to pinned host memory for the
k-points listed in host_k and are dropped for the others (G_loc, Sigma and every
observable of the loop are whole on the device).  That is what lets the full 16x16 grid
at norb 8, ntau 600, nt 1000 fit on one GPU."""
import numpy as np

from . import cntr


def peierls_A(t, E0=0.5, omega=2.0):
    return E0 * np.sin(omega * t) * np.exp(-(t - 5.0) ** 2 / 2.0)


def random_hermitian(norb, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(scale=0.5, size=(norb, norb)) + 1j * rng.normal(scale=0.5, size=(norb, norb))
    return 0.5 * (X + X.conj().T)


def hamiltonians(nkx, norb, nt, h, seed=20240601, E0=0.5, omega=2.0):
    """H(k, t) for t = -1..nt at index t+1, k = (2 pi i / nkx, 2 pi j / nkx): (Nk, nt+2, norb, norb)"""
    Hr = random_hermitian(norb, seed)
    ks = 2.0 * np.pi * np.arange(nkx) / nkx
    A = np.zeros(nt + 2)
    A[1:] = peierls_A(np.arange(nt + 1) * h, E0, omega)
    eye = np.eye(norb)
    out = np.empty((nkx * nkx, nt + 2, norb, norb), np.complex128)
    q = 0
    for kx in ks:
        for ky in ks:
            eps = -2.0 * (np.cos(kx - A) + np.cos(ky - A))
            out[q] = Hr[None, :, :] + eps[:, None, None] * eye[None, :, :]
            q += 1
    return out


class MultiOrbitalLattice:
    def __init__(self, nkx, norb, beta, ntau, nt, h, g2=0.09, SolveOrder=5, mu=0.0, seed=20240601,
                 E0=0.5, omega=2.0, window=None, host_k=()):
        import torch
        self.nk, self.norb, self.beta, self.ntau, self.nt, self.h = nkx * nkx, norb, beta, ntau, nt, h
        self.g2, self.k, self.mu = g2, SolveOrder, mu
        self.Hh = hamiltonians(nkx, norb, nt, h, seed, E0, omega)
        self.H = []
        for q in range(self.nk):
            f = cntr.function(nt, norb)
            f.data.copy_(torch.from_numpy(self.Hh[q]))
            self.H.append(f)
        if window is None:
            self.G = [cntr.herm_matrix(nt, ntau, norb) for _ in range(self.nk)]
        else:
            self.G = [cntr.herm_matrix_window(nt, ntau, norb, window=window, SolveOrder=SolveOrder,
                                              host=q in host_k) for q in range(self.nk)]
        self.window = window
        self.Sig = cntr.herm_matrix(nt, ntau, norb)
        self.Gloc = cntr.herm_matrix(nt, ntau, norb)
        self.Gold = cntr.herm_matrix(nt, ntau, norb)
        self._w = [1.0 / self.nk] * self.nk
        self.iterations = {}

    def _gloc(self, t):
        cntr.timestep_lincomb(t, self.Gold, [self.Gloc], [1.0])
        cntr.timestep_lincomb(t, self.Gloc, self.G, self._w)

    def _sigma(self, t):
        cntr.timestep_lincomb(t, self.Sig, [self.Gloc], [self.g2])

    def matsubara(self, tol=1e-12, maxiter=50, progress=None):
        for it in range(maxiter):
            if progress is not None:
                progress(it)
            self._sigma(-1)
            for q in range(self.nk):
                cntr.dyson_mat(self.G[q], self.mu, self.H[q], self.Sig, self.beta, self.k)
            self._gloc(-1)
            if it > 0 and cntr.distance_norm2(-1, self.Gloc, self.Gold) < tol:
                break
        self.iterations[-1] = it + 1
        return it + 1

    def start(self, tol=1e-12, maxiter=50):
        K = self.k
        for it in range(maxiter):
            for t in range(K + 1):
                self._sigma(t)
            for q in range(self.nk):
                cntr.dyson_start(self.G[q], self.mu, self.H[q], self.Sig, self.beta, self.h, K)
            err = 0.0
            for t in range(K + 1):
                self._gloc(t)
                err += cntr.distance_norm2(t, self.Gloc, self.Gold)
            if it > 0 and err < tol:
                break
        self.iterations[K] = it + 1
        return it + 1

    def step(self, n, tol=1e-12, maxiter=50):
        if self.window is not None:
            for g in self.G:
                g.advance(n)
        for q in range(self.nk):
            cntr.extrapolate_timestep(n - 1, self.G[q], self.k)
        cntr.timestep_lincomb(n, self.Gloc, self.G, self._w)
        for it in range(maxiter):
            self._sigma(n)
            cntr.dyson_timestep_batched(n, self.G, self.mu, self.H, [self.Sig] * self.nk, self.beta,
                                        self.h, self.k)
            self._gloc(n)
            if cntr.distance_norm2(n, self.Gloc, self.Gold) < tol:
                break
        self.iterations[n] = it + 1
        return it + 1

    def finish(self):
        """hand the resident timesteps of windowed k-points to the host"""
        if self.window is not None:
            for g in self.G:
                g.flush()
