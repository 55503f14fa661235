"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this
module; the product package nessi_amd never does.
"""
import contextlib
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")


class OrcHerm(ctypes.Structure):
    _fields_ = [("nt", ctypes.c_int), ("ntau", ctypes.c_int), ("size1", ctypes.c_int),
                ("sig", ctypes.c_int), ("ret", ctypes.c_void_p), ("les", ctypes.c_void_p),
                ("tv", ctypes.c_void_p), ("mat", ctypes.c_void_p)]


class OrcFunc(ctypes.Structure):
    _fields_ = [("nt", ctypes.c_int), ("size1", ctypes.c_int), ("data", ctypes.c_void_p)]


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def load(path):
    """a CDLL of an oracle library at `path` with every entry point's signature set"""
    so = ctypes.CDLL(path)
    P = ctypes.POINTER
    H, F = P(OrcHerm), P(OrcFunc)
    d, i, vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
    sig = {
        "orc_integrator_tables": [i, vp, vp, vp, vp, vp, vp, vp],
        "orc_gregory_weight": [i, i, i],
        "orc_convolution_matsubara": [H, H, H, d, i],
        "orc_convolution_timestep": [i, H, H, H, H, H, d, d, i],
        "orc_incr_convolution": [i, vp, d, d, H, H, H, vp, vp, H, H, d, d, i],
        "orc_dyson_mat": [H, d, F, H, d, i, i, i],
        "orc_dyson_start": [H, d, F, H, d, d, i],
        "orc_dyson_timestep": [i, H, d, F, H, d, d, i],
        "orc_vie2_mat": [H, H, H, H, d, i, i],
        "orc_vie2_start": [H, H, H, H, d, d, i],
        "orc_vie2_timestep": [i, H, H, H, H, d, d, i, i],
        "orc_bubble1": [i, H, i, i, H, H, i, i, H, H, i, i],
        "orc_bubble2": [i, H, i, i, H, H, i, i, H, H, i, i],
        "orc_extrapolate_timestep": [i, H, i],
        "orc_distance_norm2": [i, H, H],
        "orc_green_from_H_const": [H, d, vp, d, d],
        "orc_green_from_H_td": [H, d, F, d, d, i, i, i, i],
        "orc_force_matsubara_hermitian": [H],
        "orc_lattice_run": [i, vp, d, d, i, i, d, i, i, d, i, i, i, vp, vp, vp, H, i, vp, vp],
        "orc_dyson_solve": [H, H, d, F, vp, d, d, d, i, i, i],
    }
    for name, args in sig.items():
        fn = getattr(so, name)
        fn.argtypes = args
        fn.restype = ctypes.c_double if name in ("orc_gregory_weight", "orc_distance_norm2") \
            else ctypes.c_int
    return so


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = load(_LIB_PATH)
    return _lib


@contextlib.contextmanager
def using(path):
    """every oracle call inside the block goes to the library at `path` (a deliberately
    altered copy, built by the tests that check the pins catch such alterations)"""
    global _lib
    saved = lib()
    _lib = load(path)
    try:
        yield
    finally:
        _lib = saved


class HostHerm:
    """herm_matrix storage in host numpy arrays with the reference's packed layout
    (libcntr/cntr/cntr_herm_matrix_impl.hpp:62-109, 456-478)."""

    def __init__(self, nt, ntau, size1, sig=-1):
        self.nt, self.ntau, self.size1, self.sig = nt, ntau, size1, sig
        s = size1
        ntri = (nt + 1) * (nt + 2) // 2 if nt >= 0 else 0
        ntv = (nt + 1) * (ntau + 1) if nt >= 0 else 0
        self.ret = np.zeros((max(ntri, 1), s, s), np.complex128)
        self.les = np.zeros((max(ntri, 1), s, s), np.complex128)
        self.tv = np.zeros((max(ntv, 1), s, s), np.complex128)
        self.mat = np.zeros((ntau + 1, s, s), np.complex128)

    # element accessors mirroring retptr/lesptr/tvptr/matptr
    def ret_at(self, i, j):
        return self.ret[i * (i + 1) // 2 + j]

    def les_at(self, i, j):
        return self.les[j * (j + 1) // 2 + i]

    def tv_at(self, i, m):
        return self.tv[i * (self.ntau + 1) + m]

    def ret_row(self, i):
        o = i * (i + 1) // 2
        return self.ret[o:o + i + 1]

    def les_col(self, j):
        o = j * (j + 1) // 2
        return self.les[o:o + j + 1]

    def tv_row(self, i):
        return self.tv[i * (self.ntau + 1):(i + 1) * (self.ntau + 1)]

    def desc(self):
        d = OrcHerm(self.nt, self.ntau, self.size1, self.sig,
                    self.ret.ctypes.data, self.les.ctypes.data, self.tv.ctypes.data,
                    self.mat.ctypes.data)
        d._keep = self
        return d

    def copy(self):
        c = HostHerm(self.nt, self.ntau, self.size1, self.sig)
        for a in ("ret", "les", "tv", "mat"):
            getattr(c, a)[...] = getattr(self, a)
        return c


class HostFunc:
    """cntr::function storage: H(t) at index t+1, t=-1..nt."""

    def __init__(self, nt, size1):
        self.nt, self.size1 = nt, size1
        self.data = np.zeros((nt + 2, size1, size1), np.complex128)

    def __setitem__(self, t, v):
        self.data[t + 1] = v

    def __getitem__(self, t):
        return self.data[t + 1]

    def desc(self):
        d = OrcFunc(self.nt, self.size1, self.data.ctypes.data)
        d._keep = self
        return d


def _p(x):
    return ctypes.byref(x.desc())


def integrator_tables(k):
    k1 = k + 1
    out = dict(greg=np.zeros(4 * k1 * k1), omega=np.zeros(k1), bd=np.zeros(k1 + 1),
               P=np.zeros(k1 * k1), D=np.zeros(k1 * k1), W=np.zeros(k1 ** 3),
               rcorr=np.zeros(max(k - 1, 0) * k1 * k1 + 1))
    rc = lib().orc_integrator_tables(k, *[out[n].ctypes.data for n in
                                          ("greg", "omega", "bd", "P", "D", "W", "rcorr")])
    assert rc == 0
    out["rcorr"] = out["rcorr"][:max(k - 1, 0) * k1 * k1]
    return out


def gregory_weight(k, n, j):
    return lib().orc_gregory_weight(k, n, j)


def convolution_matsubara(C, A, B, beta, k=5):
    return lib().orc_convolution_matsubara(_p(C), _p(A), _p(B), beta, k)


def convolution_timestep(n, C, A, Acc, B, Bcc, beta, h, k=5):
    return lib().orc_convolution_timestep(n, _p(C), _p(A), _p(Acc), _p(B), _p(Bcc), beta, h, k)


def incr_convolution(tstp, mask, alpha, C, A, Acc, f0, ft, B, Bcc, beta, h, k=5):
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    f0p = None if f0 is None else np.ascontiguousarray(f0, np.complex128)
    ftp = None if ft is None else np.ascontiguousarray(ft, np.complex128)
    return lib().orc_incr_convolution(
        tstp, None if m is None else m.ctypes.data, complex(alpha).real, complex(alpha).imag, _p(C), _p(A), _p(Acc),
        None if f0p is None else f0p.ctypes.data, None if ftp is None else ftp.ctypes.data,
        _p(B), _p(Bcc), beta, h, k)


def dyson_mat(G, mu, H, Sigma, beta, k=5, method=2, force_hermitian=True):
    return lib().orc_dyson_mat(_p(G), mu, _p(H), _p(Sigma), beta, k, method,
                               int(force_hermitian))


def dyson_start(G, mu, H, Sigma, beta, h, k=5):
    return lib().orc_dyson_start(_p(G), mu, _p(H), _p(Sigma), beta, h, k)


def dyson_timestep(n, G, mu, H, Sigma, beta, h, k=5):
    return lib().orc_dyson_timestep(n, _p(G), mu, _p(H), _p(Sigma), beta, h, k)


def vie2_mat(G, F, Fcc, Q, beta, k=5, method=2):
    return lib().orc_vie2_mat(_p(G), _p(F), _p(Fcc), _p(Q), beta, k, method)


def vie2_start(G, F, Fcc, Q, beta, h, k=5):
    return lib().orc_vie2_start(_p(G), _p(F), _p(Fcc), _p(Q), beta, h, k)


def vie2_timestep(n, G, F, Fcc, Q, beta, h, k=5, method=2):
    return lib().orc_vie2_timestep(n, _p(G), _p(F), _p(Fcc), _p(Q), beta, h, k, method)


def bubble1(tstp, C, c1, c2, A, Acc, a1, a2, B, Bcc, b1, b2):
    return lib().orc_bubble1(tstp, _p(C), c1, c2, _p(A), _p(Acc), a1, a2, _p(B), _p(Bcc), b1, b2)


def bubble2(tstp, C, c1, c2, A, Acc, a1, a2, B, Bcc, b1, b2):
    return lib().orc_bubble2(tstp, _p(C), c1, c2, _p(A), _p(Acc), a1, a2, _p(B), _p(Bcc), b1, b2)


def extrapolate_timestep(n, G, k=5):
    return lib().orc_extrapolate_timestep(n, _p(G), k)


def distance_norm2(tstp, g1, g2):
    return lib().orc_distance_norm2(tstp, _p(g1), _p(g2))


def green_from_H_const(G, mu, H0, beta, h):
    H0 = np.ascontiguousarray(np.asarray(H0, np.complex128).reshape(G.size1, G.size1))
    return lib().orc_green_from_H_const(_p(G), mu, H0.ctypes.data, beta, h)


def green_from_H_td(G, mu, H, beta, h, SolveOrder=5, cf_order=4, fixHam=True, tstp=-2):
    """time-dependent H (cntr_equilibrium_impl.hpp:754-870); tstp >= -1: the timestep form"""
    rc = lib().orc_green_from_H_td(_p(G), mu, _p(H), beta, h, SolveOrder, cf_order, int(fixHam), tstp)
    if rc:
        raise ValueError("cf_order must be 2 or 4")


def dyson_solve(r, nt_stop=None):
    """The engine's request r (nessi_amd.engine.make_request) solved on the host: Sigma =
    V^2 g[Hbath], dyson_mat, dyson_start, dyson_timestep up to nt_stop (default nt).
    Returns (G, Sigma) as HostHerm."""
    s, nt, ntau, sig = r["size1"], r["nt"], r["ntau"], r["sig"]
    G, S = HostHerm(nt, ntau, s, sig), HostHerm(nt, ntau, s, sig)
    H = HostFunc(nt, s)
    H.data[...] = r["H"]
    Hb = np.ascontiguousarray(np.asarray(r["Hbath"], np.complex128).reshape(s, s))
    rc = lib().orc_dyson_solve(_p(G), _p(S), r["mu"], _p(H), Hb.ctypes.data, r["V"], r["beta"],
                               r["h"], r["SolveOrder"], r["mat_method"],
                               nt if nt_stop is None else nt_stop)
    assert rc == 0
    return G, S


def force_matsubara_hermitian(G):
    return lib().orc_force_matsubara_hermitian(_p(G))


def lattice_run(epsk, U, beta, ntau, nt, h, k, nt_stop, tol, maxiter, nthreads, want_gloc=False,
                setup_threads=None, kidx=()):
    """Self-consistent 2nd-order lattice on the CPU (OpenMP over k).  Returns the local
    density per timestep, the total number of corrector iterations, the seconds spent
    in the time stepping on `nthreads` threads (setup excluded, run on `setup_threads`),
    G_loc (HostHerm) if asked for, and the G_k listed in kidx (list of HostHerm)."""
    epsk = np.ascontiguousarray(epsk, np.float64)
    dens = np.zeros(nt_stop + 1)
    iters = ctypes.c_int(0)
    secs = ctypes.c_double(0)
    gl = HostHerm(nt, ntau, 1, -1) if want_gloc else None
    gks = [HostHerm(nt, ntau, 1, -1) for _ in kidx]
    kid = np.ascontiguousarray(list(kidx) or [0], np.int32)
    descs = (OrcHerm * max(len(gks), 1))(*[g.desc() for g in gks])
    rc = lib().orc_lattice_run(len(epsk), epsk.ctypes.data, U, beta, ntau, nt, h, k, nt_stop,
                               tol, maxiter, nthreads, setup_threads or nthreads, dens.ctypes.data,
                               ctypes.byref(iters), ctypes.byref(secs),
                               ctypes.byref(gl.desc()) if gl else None, len(gks), kid.ctypes.data,
                               ctypes.cast(descs, ctypes.c_void_p))
    assert rc == 0
    if kidx:
        return dens, iters.value, secs.value, gl, gks
    return dens, iters.value, secs.value, gl
