"""k-batched self-consistent lattice on one GPU (driver of the C ABI's cntrx_lattice_*).

Replaces the caller loops of examples/programs/hubbard_chain_2b.cpp:310-322 (second-order
self-energy, predictor/corrector) and the MPI k-parallel loop of
examples/programs/gw.cpp:386-430 with one device holding every k-point.
"""
import ctypes
import os

import numpy as np
import torch

from . import _lib
from ._lib import CntrxHerm, check


class Lattice:
    def __init__(self, epsk, U, beta, ntau, nt, h, SolveOrder=5, mu=0.0):
        self.nk, self.ntau, self.nt, self.k = len(epsk), ntau, nt, SolveOrder
        self.epsk = np.ascontiguousarray(epsk, np.float64)
        self._h = ctypes.c_void_p()
        check(_lib.load().cntrx_lattice_create(self.nk, self.epsk.ctypes.data, U, mu, beta, ntau,
                                               nt, h, SolveOrder, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            check(_lib.load().cntrx_lattice_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _stream():
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def equilibrium(self):
        check(_lib.load().cntrx_lattice_equilibrium(self._h, self._stream()))

    def bootstrap(self, tol=1e-12, maxiter=50):
        it = ctypes.c_int(0)
        check(_lib.load().cntrx_lattice_bootstrap(self._h, tol, maxiter, ctypes.byref(it),
                                                  self._stream()))
        return it.value

    def step(self, tstp, batched=True, tol=1e-12, maxiter=50):
        it, err = ctypes.c_int(0), ctypes.c_double(0)
        check(_lib.load().cntrx_lattice_step(self._h, tstp, int(batched), tol, maxiter,
                                             ctypes.byref(it), ctypes.byref(err), self._stream()))
        return it.value, err.value

    def run(self, t0, t1, tol=1e-12, maxiter=50):
        """Batched timesteps t0..t1 driven from the device, bit for bit the same as
        step(t, True, tol, maxiter) for t = t0..t1; returns the corrector iterations and
        the last distance of each timestep."""
        n = t1 - t0 + 1
        it = np.zeros(max(n, 1), np.int32)
        err = np.zeros(max(n, 1), np.float64)
        check(_lib.load().cntrx_lattice_run(self._h, t0, t1, tol, maxiter, it.ctypes.data,
                                            err.ctypes.data, self._stream()))
        return it[:n], err[:n]

    def fused(self):
        """True when the mixed component and the retarded row of an iteration run as one
        launch (k_tvret1); the kernel timing then reports that launch as kernel 0."""
        f = ctypes.c_int(0)
        check(_lib.load().cntrx_lattice_fused(self._h, ctypes.byref(f)))
        return bool(f.value)

    PROFILE_OFF, PROFILE_ALL, PROFILE_SAMPLED = 0, 1, 6

    def profile(self, mode=PROFILE_ALL):
        """HIP-event timing of the batched Dyson kernels: 0 off, 1 all four kernels every
        iteration, 2+q kernel q only, 6 sampled (every 8th iteration times one kernel,
        the four in turn; the other iterations run as the full graph)."""
        check(_lib.load().cntrx_lattice_profile(self._h, int(mode)))

    def kernel_stats(self):
        ms = (ctypes.c_double * 4)()
        n = (ctypes.c_long * 4)()
        b = ctypes.c_double(0)
        check(_lib.load().cntrx_lattice_kernel_stats(self._h, ms, n, ctypes.byref(b)))
        names = ("k_tv", "k_retrec", "k_lesconv", "k_lesrec")
        return {"ms": dict(zip(names, list(ms))), "launches": dict(zip(names, list(n))),
                "tv_bytes": b.value}

    def kernel_samples(self):
        """Timed launches since the profile mode was set: (kernel index, timestep, ms)."""
        import numpy as np
        lib = _lib.load()
        n = ctypes.c_long(0)
        check(lib.cntrx_lattice_kernel_samples(self._h, 0, None, None, None, ctypes.byref(n)))
        q = np.zeros(n.value, np.int32)
        t = np.zeros(n.value, np.int32)
        ms = np.zeros(n.value, np.float32)
        check(lib.cntrx_lattice_kernel_samples(self._h, n.value, q.ctypes.data, t.ctypes.data,
                                               ms.ctypes.data, ctypes.byref(n)))
        return q, t, ms

    def matrix(self, which):
        d = CntrxHerm()
        check(_lib.load().cntrx_lattice_matrix(self._h, which, ctypes.byref(d)))
        return d

    def fetch(self, which):
        """copy G_k (which >= 0), G_loc (-1) or Sigma (-2) to host numpy arrays"""
        d = self.matrix(which)
        return herm_to_host(d)

    # --- checkpoints in the reference's text format ---------------------------------
    def _ckpt_files(self, directory):
        names = [(q, "G_k%d.dat" % q) for q in range(self.nk)] + [(-1, "Gloc.dat"), (-2, "Sigma.dat")]
        return [(q, os.path.join(directory, f)) for q, f in names]

    def write_checkpoint(self, directory, precision=17):
        """Every G_k, G_loc and Sigma as herm_matrix::print_to_file text files
        (libcntr/cntr/cntr_herm_matrix_impl.hpp:1308-1359) under directory: G_k<q>.dat,
        Gloc.dat, Sigma.dat.  17 digits reproduce every value exactly."""
        lib = _lib.load()
        os.makedirs(directory, exist_ok=True)
        for q, path in self._ckpt_files(directory):
            d = self.matrix(q)
            check(lib.cntrx_herm_print_to_file(ctypes.byref(d), os.fsencode(path), int(precision),
                                               self._stream()))

    def read_checkpoint(self, directory, nt1):
        """Matsubara parts and timesteps 0..nt1 of every G_k, G_loc and Sigma from
        write_checkpoint's files (herm_matrix::read_from_file(nt1, file),
        libcntr/cntr/cntr_herm_matrix_impl.hpp:1469-1566), in place of equilibrium() and
        bootstrap() and the timesteps up to nt1; run(nt1 + 1, ...) then continues the
        propagation.  Sigma's timestep n depends only on G_loc's, so these three are the
        whole state the later timesteps read."""
        lib = _lib.load()
        for q, path in self._ckpt_files(directory):
            d = self.matrix(q)
            check(lib.cntrx_herm_read_from_file_nt1(ctypes.byref(d), int(nt1), os.fsencode(path),
                                                    self._stream()))

    def density(self, tstp, which=-1):
        a = self.fetch(which)
        if tstp == -1:
            return float((-a["mat"][-1, 0, 0]).real)
        return float((-1j * a["les"][tstp * (tstp + 1) // 2 + tstp, 0, 0]).real)


def herm_to_host(d):
    """device cntrx_herm descriptor -> dict of host arrays (cntrx_copy, synchronous)"""
    nt, ntau, s1 = d.nt, d.ntau, d.size1
    ntri = (nt + 1) * (nt + 2) // 2 if nt >= 0 else 0
    ntv = (nt + 1) * (ntau + 1) if nt >= 0 else 0
    out = {}
    for name, ptr, n in (("ret", d.ret, ntri), ("les", d.les, ntri), ("tv", d.tv, ntv),
                         ("mat", d.mat, ntau + 1)):
        buf = np.empty((max(n, 1), s1, s1), np.complex128)
        if n:
            check(_lib.load().cntrx_copy(buf.ctypes.data, ptr, n * s1 * s1 * 16, 2, None))
        out[name] = buf
    return out


def chain_1d(nk):
    """eps_k = -2 cos(k), k = 2 pi q / nk."""
    return -2.0 * np.cos(2.0 * np.pi * np.arange(nk) / nk)
