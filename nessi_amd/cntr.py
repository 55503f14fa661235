"""Python mirror of the reference's `cntr::` interface for the Kadanoff-Baym path.

Names, argument order and meaning follow libcntr (namespace cntr); every call goes
through the C ABI of libcntrx.so (include/cntrx.h) and runs on the GPU.  Device
memory is held in torch tensors (plumbing only).  There is no CPU fallback: without
the built library or a HIP device the calls raise.
"""
import ctypes
import os

import numpy as np
import torch

from . import _lib
from ._lib import (CNTRX_MAT_CG, CNTRX_MAT_FIXPOINT, CNTRX_MAT_FOURIER, CntrxFunc, CntrxHerm,
                   check)

CNTR_MAT_FOURIER = CNTRX_MAT_FOURIER
CNTR_MAT_CG = CNTRX_MAT_CG
CNTR_MAT_FIXPOINT = CNTRX_MAT_FIXPOINT
FERMION, BOSON = -1, 1
MAX_SOLVE_ORDER = 5


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _path(file):
    return os.fsencode(file)


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError("nessi_amd needs a HIP device (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


class herm_matrix:
    """cntr::herm_matrix<double> with device storage in the reference's packed layout
    (libcntr/cntr/cntr_herm_matrix_decl.hpp:44-301, _impl.hpp:62-109)."""

    def __init__(self, nt, ntau, size1=1, sig=FERMION, device=None, storage=None):
        # the solvers take size1 1..8 (3 and 5..7 run embedded in 4 and 8); the element-wise
        # utilities (bubbles, extrapolate_timestep, distance_norm2) take any size1
        if not (isinstance(size1, int) and size1 >= 1):
            raise ValueError("size1 must be a positive integer")
        self.nt, self.ntau, self.size1, self.sig = nt, ntau, size1, sig
        dev = device or _device()
        s = size1
        ntri = (nt + 1) * (nt + 2) // 2 if nt >= 0 else 0
        ntv = (nt + 1) * (ntau + 1) if nt >= 0 else 0
        if storage is None:
            storage = torch.zeros(2 * max(ntri, 1) + max(ntv, 1) + ntau + 1, s, s,
                                  dtype=torch.complex128, device=dev)
        self._storage = storage
        o = 0
        self.ret = storage[o:o + max(ntri, 1)]
        o += max(ntri, 1)
        self.les = storage[o:o + max(ntri, 1)]
        o += max(ntri, 1)
        self.tv = storage[o:o + max(ntv, 1)]
        o += max(ntv, 1)
        self.mat = storage[o:o + ntau + 1]
        self._desc = CntrxHerm(nt, ntau, size1, sig, self.ret.data_ptr(), self.les.data_ptr(),
                               self.tv.data_ptr(), self.mat.data_ptr())

    @staticmethod
    def storage_elems(nt, ntau, size1):
        ntri = (nt + 1) * (nt + 2) // 2 if nt >= 0 else 0
        ntv = (nt + 1) * (ntau + 1) if nt >= 0 else 0
        return 2 * max(ntri, 1) + max(ntv, 1) + ntau + 1

    @property
    def desc(self):
        return self._desc

    def element_size(self):
        return self.size1 * self.size1

    # --- host transfer ------------------------------------------------------------
    def to_numpy(self):
        return {c: getattr(self, c).cpu().numpy() for c in ("ret", "les", "tv", "mat")}

    def from_numpy(self, arrs):
        for c in ("ret", "les", "tv", "mat"):
            if c in arrs:
                getattr(self, c).copy_(torch.from_numpy(np.ascontiguousarray(arrs[c])))
        return self

    # --- timestep algebra (cntr_herm_matrix_impl.hpp:2064-2400) --------------------
    def set_timestep_zero(self, tstp):
        """(libcntr/cntr/cntr_herm_matrix_impl.hpp:2065-2074): mat for tstp = -1,
        else ret row, tv row and les column tstp"""
        if tstp == -1:
            self.mat.zero_()
            return
        o = tstp * (tstp + 1) // 2
        self.ret[o:o + tstp + 1].zero_()
        self.les[o:o + tstp + 1].zero_()
        self.tv[tstp * (self.ntau + 1):(tstp + 1) * (self.ntau + 1)].zero_()

    def set_timestep(self, tstp, other):
        _lincomb(tstp, self, [other], [1.0])

    def incr_timestep(self, tstp, other, alpha=1.0):
        _lincomb(tstp, self, [self, other], [1.0, float(alpha)])

    def smul(self, tstp, weight):
        _lincomb(tstp, self, [self], [float(weight)])

    def left_multiply(self, tstp, f, weight=1.0):
        check(_lib.load().cntrx_timestep_scale_by_function(tstp, self.desc, f.host_desc(),
                                                           float(weight), 1, _stream()))

    def right_multiply(self, tstp, f, weight=1.0):
        check(_lib.load().cntrx_timestep_scale_by_function(tstp, self.desc, f.host_desc(),
                                                           float(weight), 0, _stream()))

    # --- text checkpoints (cntr_herm_matrix_impl.hpp:1308-1566) ----------------------
    def print_to_file(self, file, precision=16):
        """herm_matrix::print_to_file (libcntr/cntr/cntr_herm_matrix_impl.hpp:1308-1359):
        the reference's text format; precision 17 reproduces every value exactly"""
        check(_lib.load().cntrx_herm_print_to_file(self.desc, _path(file), int(precision),
                                                   _stream()))

    def read_from_file(self, *args):
        """read_from_file(file) or read_from_file(nt1, file)
        (libcntr/cntr/cntr_herm_matrix_impl.hpp:1376-1448, 1469-1566).  The first form
        reallocates G where the reference resizes it (a file with more timesteps, another
        ntau or size1) and takes the file's sig; the second reads mat and timesteps 0..nt1
        and needs nt1 <= nt, nt1 <= the file's nt, the same ntau and size1."""
        lib = _lib.load()
        if len(args) == 1:
            file = _path(args[0])
            n, m, s1, sg = (ctypes.c_int() for _ in range(4))
            check(lib.cntrx_herm_file_header(file, ctypes.byref(n), ctypes.byref(m),
                                             ctypes.byref(s1), ctypes.byref(sg)))
            if n.value > self.nt or m.value != self.ntau or s1.value != self.size1:
                self.__init__(n.value, m.value, s1.value, self.sig, device=self._storage.device)
            sig = ctypes.c_int()
            check(lib.cntrx_herm_read_from_file(self.desc, file, ctypes.byref(sig), _stream()))
            self.sig = self._desc.sig = sig.value
        elif len(args) == 2:
            check(lib.cntrx_herm_read_from_file_nt1(self.desc, int(args[0]), _path(args[1]),
                                                    _stream()))
        else:
            raise TypeError("read_from_file(file) or read_from_file(nt1, file)")

    def density_matrix(self, tstp):
        """(libcntr/cntr/cntr_herm_matrix_impl.hpp:983-1028)"""
        if tstp == -1:
            return -self.mat[self.ntau].cpu().numpy()
        o = tstp * (tstp + 1) // 2 + tstp
        return (1j * self.sig) * self.les[o].cpu().numpy()


_desc_generation = [0]  # bumped whenever a descriptor's pointers move


def _tri(n):
    return n * (n + 1) // 2


class herm_matrix_window(herm_matrix):
    """A herm_matrix whose retarded rows and lesser columns are resident on the device only
    in a window of `window` timesteps; the mixed and Matsubara components stay whole.

    A Dyson timestep n reads and writes G's retarded row n, lesser column n and mixed rows
    0..n only (libcntr/cntr/cntr_dyson_impl.hpp:79-223, 650-711, 892-1022; the history it
    integrates over is Sigma's), and the predictor extrapolate_timestep(n-1) reads the
    rows n-1-k..n-1 (cntr_utilities_impl.hpp:21-170).  So a long propagation needs the
    whole G^tv history per k-point (the largest component at ntau ~ nt) but only the last
    k+2 retarded rows and lesser columns.  Rows that leave the window go to pinned host
    memory (host=True) or are dropped.

    The window buffers hold consecutive rows of the packed triangles (row i of ret at
    i(i+1)/2, column j of les at j(j+1)/2, as in herm_matrix); the descriptor's ret and les
    pointers are rebased by the offset of the window's first row, so every kernel addresses
    a resident row exactly as in a full herm_matrix.  Call advance(n) before extrapolating
    into or solving timestep n; call flush() after the last timestep."""

    def __init__(self, nt, ntau, size1=1, sig=FERMION, window=16, SolveOrder=MAX_SOLVE_ORDER,
                 host=False, device=None):
        # a call at size1 3 or 5..7 copies timesteps 0..n of its arguments, which a window
        # does not hold
        if size1 not in (1, 2, 4, 8):
            raise ValueError("a windowed herm_matrix needs size1 1, 2, 4 or 8")
        if window < 2 * (SolveOrder + 2):
            raise ValueError("window must hold at least 2 (SolveOrder + 2) timesteps")
        self.nt, self.ntau, self.size1, self.sig = nt, ntau, size1, sig
        self.window, self.k = window, SolveOrder
        dev = device or _device()
        s = size1
        cap = window * (nt + 1)  # entries of `window` rows of at most nt+1 entries
        ntv = (nt + 1) * (ntau + 1)
        self._storage = torch.zeros(ntv + ntau + 1 + 2 * cap, s, s, dtype=torch.complex128, device=dev)
        self.tv = self._storage[0:ntv]
        self.mat = self._storage[ntv:ntv + ntau + 1]
        self.ret_window = self._storage[ntv + ntau + 1:ntv + ntau + 1 + cap]
        self.les_window = self._storage[ntv + ntau + 1 + cap:]
        self.host = host
        if host:
            ntri = _tri(nt + 1)
            self.ret = torch.zeros(ntri, s, s, dtype=torch.complex128, pin_memory=True)
            self.les = torch.zeros(ntri, s, s, dtype=torch.complex128, pin_memory=True)
        else:
            self.ret = self.les = None
        self.r0 = 0  # first resident timestep
        self.streamed = 0  # timesteps handed to the host or dropped
        self._rebase()

    def _rebase(self):
        es_bytes = self.size1 * self.size1 * 16
        off = _tri(self.r0) * es_bytes
        self._desc = CntrxHerm(self.nt, self.ntau, self.size1, self.sig,
                               self.ret_window.data_ptr() - off, self.les_window.data_ptr() - off,
                               self.tv.data_ptr(), self.mat.data_ptr())
        _desc_generation[0] += 1

    def resident(self):
        """first and last timestep whose retarded row and lesser column are on the device"""
        return self.r0, min(self.r0 + self.window - 1, self.nt)

    def _stream_out(self, r1):
        """hand timesteps r0..r1-1 to the host (or drop them)"""
        if r1 <= self.r0:
            return
        a, b = _tri(self.r0), _tri(r1)
        if self.host:
            self.ret[a:b].copy_(self.ret_window[0:b - a], non_blocking=True)
            self.les[a:b].copy_(self.les_window[0:b - a], non_blocking=True)
        self.streamed = r1

    def advance(self, n):
        """make timestep n resident, keeping the k+1 timesteps before it"""
        if n < 0 or n > self.nt:
            raise ValueError("timestep out of range")
        if n < self.r0 + self.window:
            return
        new_r0 = n - self.k - 1
        self._stream_out(new_r0)
        # move timesteps new_r0..n-1 to the front of the window buffers
        a, b, c = _tri(self.r0), _tri(new_r0), _tri(n)
        for w in (self.ret_window, self.les_window):
            w[0:c - b].copy_(w[b - a:c - a].clone())
        self.r0 = new_r0
        self._rebase()

    def flush(self, last=None):
        """hand every resident timestep up to `last` (default nt) to the host"""
        last = self.nt if last is None else last
        if self.host and last >= self.r0:
            a, b = _tri(self.r0), _tri(last + 1)
            self.ret[a:b].copy_(self.ret_window[0:b - a], non_blocking=True)
            self.les[a:b].copy_(self.les_window[0:b - a], non_blocking=True)
        self.streamed = last + 1

    def to_numpy(self):
        torch.cuda.current_stream().synchronize()
        out = {"tv": self.tv.cpu().numpy(), "mat": self.mat.cpu().numpy()}
        if self.host:
            out["ret"], out["les"] = self.ret.numpy().copy(), self.les.numpy().copy()
        return out

    def set_timestep_zero(self, tstp):
        raise NotImplementedError("herm_matrix_window: timestep algebra runs on the resident rows "
                                  "through timestep_lincomb")

    def density_matrix(self, tstp):
        if tstp == -1:
            return -self.mat[self.ntau].cpu().numpy()
        lo, hi = self.resident()
        if not lo <= tstp <= hi:
            raise ValueError("timestep %d is not resident (%d..%d)" % (tstp, lo, hi))
        o = _tri(tstp) - _tri(self.r0) + tstp
        return (1j * self.sig) * self.les_window[o].cpu().numpy()


class function:
    """cntr::function<double>: H(t), t=-1..nt, stored at index t+1
    (libcntr/cntr/cntr_function_decl.hpp:26-107)."""

    def __init__(self, nt, size1=1, device=None):
        self.nt, self.size1 = nt, size1
        self.data = torch.zeros(nt + 2, size1, size1, dtype=torch.complex128,
                                device=device or _device())
        self._desc = CntrxFunc(nt, size1, self.data.data_ptr())
        self._host = None

    @property
    def desc(self):
        return self._desc

    def set_value(self, t, value):
        self.data[t + 1] = torch.as_tensor(np.asarray(value, np.complex128).reshape(self.size1, self.size1))
        self._host = None

    def set_constant(self, value):
        v = torch.as_tensor(np.asarray(value, np.complex128).reshape(self.size1, self.size1))
        self.data[:] = v.to(self.data.device)
        self._host = None

    def print_to_file(self, file, precision=16):
        """function::print_to_file (libcntr/cntr/cntr_function_impl.hpp:719-732)"""
        check(_lib.load().cntrx_function_print_to_file(self.desc, _path(file), int(precision),
                                                       _stream()))

    def read_from_file(self, *args):
        """read_from_file(file) or read_from_file(nt1, file)
        (libcntr/cntr/cntr_function_impl.hpp:750-777, 795-824); the first form reallocates
        where the reference resizes (a file with more timesteps or another size1)"""
        lib = _lib.load()
        if len(args) == 1:
            file = _path(args[0])
            n, s1 = ctypes.c_int(), ctypes.c_int()
            check(lib.cntrx_function_file_header(file, ctypes.byref(n), ctypes.byref(s1)))
            if n.value > self.nt or s1.value != self.size1:
                self.__init__(n.value, s1.value, device=self.data.device)
            check(lib.cntrx_function_read_from_file(self.desc, file, _stream()))
        elif len(args) == 2:
            check(lib.cntrx_function_read_from_file_nt1(self.desc, int(args[0]), _path(args[1]),
                                                        _stream()))
        else:
            raise TypeError("read_from_file(file) or read_from_file(nt1, file)")
        self._host = None

    def host_desc(self):
        self._host = np.ascontiguousarray(self.data.cpu().numpy())
        d = CntrxFunc(self.nt, self.size1, self._host.ctypes.data)
        return d


def _lincomb(tstp, C, Xs, ws):
    n = len(Xs)
    arr = (CntrxHerm * n)(*[X.desc for X in Xs])
    w = (ctypes.c_double * n)(*ws)
    check(_lib.load().cntrx_timestep_lincomb(tstp, C.desc, n, ctypes.cast(arr, ctypes.c_void_p),
                                             ctypes.cast(w, ctypes.c_void_p), _stream()))


def timestep_lincomb(tstp, C, Xs, ws):
    """C(tstp) = sum_i ws[i] Xs[i](tstp) (G_loc = (1/Nk) sum_k G_k, get_loc in
    examples/programs/gw_selfene_impl.cpp:54-62)."""
    _lincomb(tstp, C, Xs, ws)


# ----------------------------------------------------------------------------------
# Dyson / VIE2 / convolution (libcntr/cntr/cntr_dyson_decl.hpp, cntr_vie2_decl.hpp,
# cntr_convolution_decl.hpp)
# ----------------------------------------------------------------------------------
def dyson_mat(G, mu, H, Sigma, beta, SolveOrder=MAX_SOLVE_ORDER, method=CNTR_MAT_FIXPOINT,
              force_hermitian=True):
    check(_lib.load().cntrx_dyson_mat(G.desc, mu, H.desc, Sigma.desc, beta, SolveOrder, method,
                                      int(force_hermitian), _stream()))


def dyson_start(G, mu, H, Sigma, beta, h, SolveOrder=MAX_SOLVE_ORDER):
    check(_lib.load().cntrx_dyson_start(G.desc, mu, H.desc, Sigma.desc, beta, h, SolveOrder,
                                        _stream()))


def dyson_timestep(n, G, mu, H, Sigma, beta, h, SolveOrder=MAX_SOLVE_ORDER):
    check(_lib.load().cntrx_dyson_timestep(n, G.desc, mu, H.desc, Sigma.desc, beta, h,
                                           SolveOrder, _stream()))


def dyson(G, mu, H, Sigma, beta, h, SolveOrder=MAX_SOLVE_ORDER, matsubara_method=CNTR_MAT_FIXPOINT,
          force_hermitian=True):
    """cntr::dyson (libcntr/cntr/cntr_dyson_impl.hpp:1725-1734)"""
    dyson_mat(G, mu, H, Sigma, beta, SolveOrder, matsubara_method, force_hermitian)
    if G.nt >= 0:
        dyson_start(G, mu, H, Sigma, beta, h, SolveOrder)
    for n in range(SolveOrder + 1, G.nt + 1):
        dyson_timestep(n, G, mu, H, Sigma, beta, h, SolveOrder)


class _Batch:
    def __init__(self, Gs, Hs, Sigmas):
        n = len(Gs)
        self.g = (CntrxHerm * n)(*[g.desc for g in Gs])
        self.s = (CntrxHerm * n)(*[s.desc for s in Sigmas])
        self.h = (CntrxFunc * n)(*[f.desc for f in Hs])


def dyson_timestep_batched(n, Gs, mu, Hs, Sigmas, beta, h, SolveOrder=MAX_SOLVE_ORDER, _cache={}):
    """All solves advance one timestep in one launch sequence (no reference
    counterpart; replaces the per-k loop of examples/programs/gw.cpp:401-411)."""
    cnt = len(Gs)
    ns = (ctypes.c_int * cnt)(*(n if hasattr(n, "__len__") else [n] * cnt))
    mus = (ctypes.c_double * cnt)(*(mu if hasattr(mu, "__len__") else [mu] * cnt))
    bs = (ctypes.c_double * cnt)(*(beta if hasattr(beta, "__len__") else [beta] * cnt))
    hs = (ctypes.c_double * cnt)(*(h if hasattr(h, "__len__") else [h] * cnt))
    key = (id(Gs), id(Hs), id(Sigmas), cnt, _desc_generation[0])
    b = _cache.get(key)
    if b is None or b[0] is not Gs:
        b = (Gs, _Batch(Gs, Hs, Sigmas))
        _cache.clear()
        _cache[key] = b
    B = b[1]
    vp = ctypes.c_void_p
    check(_lib.load().cntrx_dyson_timestep_batched(
        cnt, ctypes.cast(ns, vp), ctypes.cast(B.g, vp), ctypes.cast(mus, vp), ctypes.cast(B.h, vp),
        ctypes.cast(B.s, vp), ctypes.cast(bs, vp), ctypes.cast(hs, vp), SolveOrder, _stream()))


def vie2_mat(G, F, Fcc, Q, beta, SolveOrder=MAX_SOLVE_ORDER, method=CNTR_MAT_FIXPOINT):
    check(_lib.load().cntrx_vie2_mat(G.desc, F.desc, Fcc.desc, Q.desc, beta, SolveOrder, method,
                                     _stream()))


def vie2_start(G, F, Fcc, Q, beta, h, SolveOrder=MAX_SOLVE_ORDER):
    check(_lib.load().cntrx_vie2_start(G.desc, F.desc, Fcc.desc, Q.desc, beta, h, SolveOrder,
                                       _stream()))


def vie2_timestep(n, G, F, Fcc, Q, beta, h, SolveOrder=MAX_SOLVE_ORDER,
                  matsubara_method=CNTR_MAT_FIXPOINT):
    check(_lib.load().cntrx_vie2_timestep(n, G.desc, F.desc, Fcc.desc, Q.desc, beta, h,
                                          SolveOrder, matsubara_method, _stream()))


def convolution_timestep(n, C, A, *args):
    """convolution_timestep(n, C, A, Acc, B, Bcc, beta, h, SolveOrder) or
    convolution_timestep(n, C, A, B, beta, h, SolveOrder)
    (libcntr/cntr/cntr_convolution_decl.hpp:61-66)."""
    if isinstance(args[1], herm_matrix):
        Acc, B, Bcc, beta, h = args[:5]
        k = args[5] if len(args) > 5 else MAX_SOLVE_ORDER
    else:
        B, beta, h = args[:3]
        Acc, Bcc = A, B
        k = args[3] if len(args) > 3 else MAX_SOLVE_ORDER
    check(_lib.load().cntrx_convolution_timestep(n, C.desc, A.desc, Acc.desc, B.desc, Bcc.desc,
                                                 beta, h, k, _stream()))


def incr_convolution(tstp, alpha, C, A, Acc, f0, ft, B, Bcc, beta, h, SolveOrder=MAX_SOLVE_ORDER,
                     mask=None):
    """C += alpha (A*f*B) at timestep tstp (libcntr/cntr/cntr_convolution_decl.hpp:297-318,
    cntr_convolution_impl.hpp:4163-4805).  f0 = f(-i beta) (one size1 x size1 block) and ft =
    f(t) from t = 0 (max(tstp, SolveOrder)+1 blocks) are device tensors, or None for f = 1.
    mask: None (every output, as the public incr_convolution) or one byte per output:
    mat (ntau+1) for tstp = -1, else ret (tstp+1), tv (ntau+1), les (tstp+1) in that order."""
    alpha = complex(alpha)
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    p0 = None if f0 is None else f0.data_ptr()
    pt = None if ft is None else ft.data_ptr()
    check(_lib.load().cntrx_incr_convolution(tstp, None if m is None else m.ctypes.data,
                                             alpha.real, alpha.imag, C.desc, A.desc, Acc.desc,
                                             p0, pt, B.desc, Bcc.desc, beta, h, SolveOrder, _stream()))


def convolution_timestep_new(tstp, C, A, Acc, *args):
    """convolution_timestep_new(tstp, C, A, Acc, ft, B, Bcc, beta, h, SolveOrder) with a
    function ft, or (tstp, C, A, Acc, B, Bcc, beta, h, SolveOrder) without
    (libcntr/cntr/cntr_convolution_impl.hpp:4810-4983): C(tstp) = 0, then
    incr_convolution with alpha = 1."""
    ft = None
    if isinstance(args[0], function):
        ft, args = args[0], args[1:]
    B, Bcc, beta, h = args[:4]
    k = args[4] if len(args) > 4 else MAX_SOLVE_ORDER
    C.set_timestep_zero(tstp)
    f0 = fts = None
    if ft is not None:
        f0 = ft.data[0]
        fts = ft.data[0] if tstp == -1 else ft.data[1]
    incr_convolution(tstp, 1.0, C, A, Acc, f0, fts, B, Bcc, beta, h, k)


def convolution_matsubara(C, A, B, beta, SolveOrder=MAX_SOLVE_ORDER):
    convolution_timestep(-1, C, A, A, B, B, beta, 1.0, SolveOrder)


def _bubble(which, tstp, C, args):
    # (C, c1, c2, A, Acc, a1, a2, B, Bcc, b1, b2) or shorter reference overloads
    if len(args) == 10:
        c1, c2, A, Acc, a1, a2, B, Bcc, b1, b2 = args
    elif len(args) == 8:
        c1, c2, A, a1, a2, B, b1, b2 = args
        Acc, Bcc = A, B
    elif len(args) == 4:
        A, Acc, B, Bcc = args
        c1 = c2 = a1 = a2 = b1 = b2 = 0
    elif len(args) == 2:
        A, B = args
        Acc, Bcc = A, B
        c1 = c2 = a1 = a2 = b1 = b2 = 0
    else:
        raise TypeError("unsupported Bubble overload")
    fn = _lib.load().cntrx_bubble1 if which == 1 else _lib.load().cntrx_bubble2
    check(fn(tstp, C.desc, c1, c2, A.desc, Acc.desc, a1, a2, B.desc, Bcc.desc, b1, b2, _stream()))


def Bubble1(tstp, C, *args):
    """C_{c1c2}(t,t') = i A_{a1a2}(t,t') B_{b2b1}(t',t) (libcntr/cntr/cntr_bubble_decl.hpp)"""
    _bubble(1, tstp, C, args)


def Bubble2(tstp, C, *args):
    """C_{c1c2}(t,t') = i A_{a1a2}(t,t') B_{b1b2}(t,t') (libcntr/cntr/cntr_bubble_decl.hpp)"""
    _bubble(2, tstp, C, args)


def extrapolate_timestep(n, G, SolveOrder=MAX_SOLVE_ORDER):
    check(_lib.load().cntrx_extrapolate_timestep(n, G.desc, SolveOrder, _stream()))


def distance_norm2(tstp, g1, g2):
    out = ctypes.c_double(0.0)
    check(_lib.load().cntrx_distance_norm2(tstp, g1.desc, g2.desc, ctypes.byref(out), _stream()))
    return out.value


CNTRX_ALL_TIMESTEPS = -2


def green_from_H(*args):
    """Free GF of a quadratic Hamiltonian (libcntr/cntr/cntr_equilibrium_decl.hpp:73-98):
      green_from_H(G, mu, H0, beta, h)                         constant H0 (a matrix)
      green_from_H(G, mu, H, beta, h, SolveOrder=5, cf_order=4) H(t) a function
      green_from_H(tstp, G, mu, H, beta, h, fixHam=False, SolveOrder=5, cf_order=4)
                                                                timestep tstp only
    With a function the propagators are commutator-free exponentials of order cf_order (2 or
    4) with H interpolated at order SolveOrder (cntr_equilibrium_impl.hpp:754-870, 955-1040);
    without fixHam, H(t > SolveOrder) is replaced by its extrapolation, in H."""
    if isinstance(args[0], (int, np.integer)):
        tstp, G, mu, H, beta, h = args[:6]
        rest = list(args[6:]) + [None] * 3
        fixHam = False if rest[0] is None else bool(rest[0])
        k = MAX_SOLVE_ORDER if rest[1] is None else int(rest[1])
        cf = 4 if rest[2] is None else int(rest[2])
        return _green_td(G, mu, H, beta, h, k, cf, fixHam, int(tstp))
    G, mu, H, beta, h = args[:5]
    if isinstance(H, function):
        rest = list(args[5:]) + [None] * 2
        k = MAX_SOLVE_ORDER if rest[0] is None else int(rest[0])
        cf = 4 if rest[1] is None else int(rest[1])
        return _green_td(G, mu, H, beta, h, k, cf, True, CNTRX_ALL_TIMESTEPS)
    H0 = np.ascontiguousarray(np.asarray(H, np.complex128).reshape(G.size1, G.size1))
    check(_lib.load().cntrx_green_from_H(G.desc, mu, H0.ctypes.data, beta, h, _stream()))


def _green_td(G, mu, H, beta, h, SolveOrder, cf_order, fixHam, tstp):
    d = H.host_desc()  # the propagators are formed on the host from a host copy of H
    check(_lib.load().cntrx_green_from_H_td(G.desc, mu, ctypes.byref(d), beta, h, SolveOrder, cf_order,
                                            int(fixHam), tstp, _stream()))
    if not fixHam:  # H(t > SolveOrder) was extrapolated in place, as the reference does
        H.data.copy_(torch.from_numpy(H._host))


def force_matsubara_hermitian(G):
    check(_lib.load().cntrx_force_matsubara_hermitian(G.desc, _stream()))


def integrator_tables(k):
    k1 = k + 1
    out = dict(greg=np.zeros(4 * k1 * k1), omega=np.zeros(k1), bd=np.zeros(k1 + 1),
               P=np.zeros(k1 * k1), D=np.zeros(k1 * k1), W=np.zeros(k1 ** 3),
               rcorr=np.zeros(max(k - 1, 0) * k1 * k1 + 1))
    check(_lib.load().cntrx_integrator_tables(k, *[out[n].ctypes.data for n in
                                                   ("greg", "omega", "bd", "P", "D", "W", "rcorr")]))
    out["rcorr"] = out["rcorr"][:max(k - 1, 0) * k1 * k1]
    return out
