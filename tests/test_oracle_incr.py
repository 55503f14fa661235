"""Pins the oracle's incr_convolution (libcntr/cntr/cntr_convolution_impl.hpp:4163-4805).

1. With f = 1, alpha = 1 and C = 0 it must equal convolution_timestep
   (cntr_convolution_impl.hpp:1581-1771), which the single-pole identity and the
   convergence tests pin, on every component of every timestep (the two are different
   restatements of the same Gregory quadrature; they agree when Acc = A and B = Bcc with
   an anti-Hermitian G^<(t,t)).
2. A*f*B equals A'*B with A'(t,t') = A(t,t') f(t') and Acc'(t,t') = f(t)^dagger Acc(t,t'),
   A'^tv = A^tv f0, A'^M = A^M f0: this fixes on which side and at which time f enters
   every term (f and the blocks are random 2x2 matrices, so a transposed operand fails).
3. Masked-out outputs are untouched and the selected ones move by alpha times the
   unmasked increment."""
import numpy as np
import pytest


def _rand(orc, rng, nt, ntau, s, sig=-1, consistent=True):
    H = orc.HostHerm(nt, ntau, s, sig)
    for c in ("ret", "les", "tv", "mat"):
        a = getattr(H, c)
        a[...] = rng.normal(size=a.shape) + 1j * rng.normal(size=a.shape)
    if consistent:
        for j in range(nt + 1):
            d = H.les_at(j, j)
            d[...] = 0.5 * (d - d.conj().T)
    return H


def _step(H, t):
    if t == -1:
        return {"mat": H.mat.copy()}
    return {"ret": H.ret_row(t).copy(), "tv": H.tv_row(t).copy(), "les": H.les_col(t).copy()}


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("s,k,sig", [(1, 1, -1), (1, 5, -1), (2, 3, 1), (2, 5, -1), (3, 2, -1)])
def test_incr_without_f_is_convolution_timestep(orc, s, k, sig):
    rng = np.random.default_rng(100 + 10 * s + k)
    nt, ntau, beta, h = 18, 14, 3.0, 0.1
    A, B = _rand(orc, rng, nt, ntau, s, sig), _rand(orc, rng, nt, ntau, s, sig)
    for t in range(-1, nt + 1):
        C1, C2 = orc.HostHerm(nt, ntau, s, sig), orc.HostHerm(nt, ntau, s, sig)
        if t == -1:
            orc.convolution_matsubara(C1, A, B, beta, k)
        else:
            orc.convolution_timestep(t, C1, A, A, B, B, beta, h, k)
        orc.incr_convolution(t, None, 1.0, C2, A, A, None, None, B, B, beta, h, k)
        x1, x2 = _step(C1, t), _step(C2, t)
        for c in x1:
            assert _rel(x2[c], x1[c]) < 1e-13, (t, c)


def _fold_f(orc, A, Acc, f0, ft):
    """A'(t,t') = A(t,t') f(t'), Acc'(t,t') = f(t)^dagger Acc(t,t'), A'^tv = A^tv f0"""
    nt, ntau, s = A.nt, A.ntau, A.size1
    Ap, Accp = orc.HostHerm(nt, ntau, s, A.sig), orc.HostHerm(nt, ntau, s, A.sig)
    for i in range(nt + 1):
        for j in range(i + 1):
            Ap.ret_at(i, j)[...] = A.ret_at(i, j) @ ft[j]
            Accp.ret_at(i, j)[...] = ft[i].conj().T @ Acc.ret_at(i, j)
            Ap.les_at(j, i)[...] = A.les_at(j, i) @ ft[i]
            Accp.les_at(j, i)[...] = ft[j].conj().T @ Acc.les_at(j, i)
        for m in range(ntau + 1):
            Ap.tv_at(i, m)[...] = A.tv_at(i, m) @ f0
            Accp.tv_at(i, m)[...] = Acc.tv_at(i, m)
    for m in range(ntau + 1):
        Ap.mat[m] = A.mat[m] @ f0
        Accp.mat[m] = Acc.mat[m]
    return Ap, Accp


@pytest.mark.parametrize("k", [2, 5])
def test_incr_f_enters_as_right_factor_of_A(orc, k):
    rng = np.random.default_rng(7 + k)
    nt, ntau, s, beta, h = 16, 12, 2, 2.5, 0.05
    A, Acc = _rand(orc, rng, nt, ntau, s, consistent=False), _rand(orc, rng, nt, ntau, s, consistent=False)
    B, Bcc = _rand(orc, rng, nt, ntau, s, consistent=False), _rand(orc, rng, nt, ntau, s, consistent=False)
    f0 = rng.normal(size=(s, s)) + 1j * rng.normal(size=(s, s))
    ft = rng.normal(size=(nt + 1, s, s)) + 1j * rng.normal(size=(nt + 1, s, s))
    Ap, Accp = _fold_f(orc, A, Acc, f0, ft)
    for t in [-1, 0, 1, k - 1, k, 2 * k + 2, 2 * k + 3, nt]:
        C1, C2 = orc.HostHerm(nt, ntau, s), orc.HostHerm(nt, ntau, s)
        ftt = ft if t >= 0 else None
        orc.incr_convolution(t, None, 1.0, C1, A, Acc, f0, ftt, B, Bcc, beta, h, k)
        orc.incr_convolution(t, None, 1.0, C2, Ap, Accp, None, None, B, Bcc, beta, h, k)
        x1, x2 = _step(C1, t), _step(C2, t)
        for c in x1:
            assert _rel(x1[c], x2[c]) < 1e-13, (t, c)


def test_incr_mask_and_alpha(orc):
    rng = np.random.default_rng(3)
    nt, ntau, s, k, beta, h = 14, 10, 2, 4, 4.0, 0.1
    A, Acc = _rand(orc, rng, nt, ntau, s, consistent=False), _rand(orc, rng, nt, ntau, s, consistent=False)
    B, Bcc = _rand(orc, rng, nt, ntau, s, consistent=False), _rand(orc, rng, nt, ntau, s, consistent=False)
    alpha = 0.3 - 0.7j
    for t in [-1, 2, 11]:
        X = orc.HostHerm(nt, ntau, s)
        orc.incr_convolution(t, None, 1.0, X, A, Acc, None, None, B, Bcc, beta, h, k)
        C = _rand(orc, rng, nt, ntau, s)
        C0 = _step(C, t)
        nsel = ntau + 1 if t == -1 else 2 * (t + 1) + ntau + 1
        mask = (rng.random(nsel) < 0.5).astype(np.uint8)
        orc.incr_convolution(t, mask, alpha, C, A, Acc, None, None, B, Bcc, beta, h, k)
        got, inc = _step(C, t), _step(X, t)
        off = 0
        for c in got:
            sel = mask[off:off + len(got[c])].astype(bool)
            off += len(got[c])
            want = C0[c] + alpha * inc[c] * sel[:, None, None]
            assert np.abs(got[c] - want).max() < 1e-13 * max(1.0, np.abs(want).max()), (t, c)
