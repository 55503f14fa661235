"""The GPU convolution_timestep against closed-form values, not only against the oracle:
every component of A and B a polynomial of degree <= 2 in each argument, so each integrand
has degree <= 4 and the order-5 quadrature is exact (tests/test_oracle_convolution_exact.py
derives the fields and integrals).  size1 2, 3 (embedded in 4) and 8, fermions and bosons,
every timestep from the Matsubara branch through the start rules to n = 14: bar 1e-13
relative to the largest entry per component."""
import numpy as np
import pytest

from tests import test_oracle_convolution_exact as CX
from tests.gpu_util import to_gpu, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cntr():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nessi_amd import cntr as c
    from nessi_amd import _lib
    _lib.check(_lib.load().cntrx_init(0))
    return c


@pytest.mark.parametrize("s,sig", [(2, -1), (2, 1), (3, -1), (8, 1)])
def test_gpu_convolution_exact_for_polynomials(cntr, orc, s, sig):
    rng = np.random.default_rng(900 + 10 * s + sig)
    t, tau, P, A, Acc, B, Bcc = CX.build(orc, rng, sig, s)
    Ag, Accg, Bg, Bccg = (to_gpu(cntr, X) for X in (A, Acc, B, Bcc))
    C = cntr.herm_matrix(CX.NT, CX.NTAU, s, sig)
    for n in range(-1, CX.NT + 1):
        cntr.convolution_timestep(n, C, Ag, Accg, Bg, Bccg, CX.BETA, CX.H, CX.K)
    X = CX.expected(t, tau, P, sig)
    Ch = to_host(orc, C)
    for c in ("ret", "les", "tv", "mat"):
        err = np.abs(getattr(Ch, c) - X[c]).max() / np.abs(X[c]).max()
        assert err < 1e-13, (c, err)
