"""The conjugate-gradient Matsubara solver (CNTR_MAT_CG: cntr_vie2_impl.hpp:485-602,
cntr_dyson_impl.hpp:522-540) through the C ABI against the oracle's restatement:
dyson_mat for fermions and bosons at size1 1, 2, 3 (embedded), 4 and 8, vie2_mat with
F != Fcc at size1 2, the reference's own Matsubara test (libcntr/test/matsubara.cpp), and
the largest ntau the Matsubara solvers take.  CG stops on a residual test, so GPU and
oracle agree to the solver's tolerance rather than to rounding: bar 1e-11 relative."""
import numpy as np
import pytest

from tests import physics
from tests.gpu_util import func_to_gpu, to_gpu, to_host

pytestmark = pytest.mark.gpu
TOL = 1e-11
CG = 1


@pytest.fixture(scope="module")
def cntr():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nessi_amd import cntr as c
    from nessi_amd import _lib
    _lib.check(_lib.load().cntrx_init(0))
    return c


def rand_herm(rng, s, scale=0.5):
    X = rng.normal(scale=scale, size=(s, s)) + 1j * rng.normal(scale=scale, size=(s, s))
    return 0.5 * (X + X.conj().T)


def mat_err(orc, G, X):
    a = to_host(orc, G).mat
    return np.abs(a - X.mat).max() / np.abs(X.mat).max()


@pytest.mark.parametrize("s,sig,ntau", [(1, -1, 200), (2, -1, 200), (3, -1, 120), (4, 1, 200), (8, -1, 100),
                                         (1, -1, 4096), (4, -1, 800)])
def test_dyson_mat_cg(cntr, orc, s, sig, ntau):
    rng = np.random.default_rng(1000 + 10 * s + ntau)
    beta, mu = 6.0, 0.0
    H0, Hb = rand_herm(rng, s), rand_herm(rng, s)
    if sig == 1:  # bosons: every level above mu
        H0 = H0 + (0.6 - np.linalg.eigvalsh(H0).min()) * np.eye(s)
        Hb = Hb + (0.4 - np.linalg.eigvalsh(Hb).min()) * np.eye(s)
    H = orc.HostFunc(-1, s)
    H[-1] = H0
    Sig = orc.HostHerm(-1, ntau, s, sig)
    orc.green_from_H_const(Sig, mu, Hb, beta, 1.0)
    Sig.mat[...] *= 0.2
    X = orc.HostHerm(-1, ntau, s, sig)
    assert orc.dyson_mat(X, mu, H, Sig, beta, 5, CG, True) == 0
    G = cntr.herm_matrix(-1, ntau, s, sig)
    cntr.dyson_mat(G, mu, func_to_gpu(cntr, H), to_gpu(cntr, Sig), beta, 5, cntr.CNTR_MAT_CG)
    assert mat_err(orc, G, X) <= TOL


def test_vie2_mat_cg_with_distinct_fcc(cntr, orc):
    from tests.test_oracle_matrix_pins import HS, _embedded, _herm
    ntau, beta = 300, 5.0
    Gx, Sig, _, mu = _embedded(orc, 4, ntau, 0.05, beta)
    S = _herm(orc, -1, ntau, 2)
    S.mat[...] = Sig.mat
    G0 = physics.free_matrix(_herm(orc, -1, ntau, 2), HS, mu, beta, 0.05)
    F, Fcc = _herm(orc, -1, ntau, 2), _herm(orc, -1, ntau, 2)
    orc.convolution_matsubara(F, G0, S, beta, 5)
    orc.convolution_matsubara(Fcc, S, G0, beta, 5)
    F.mat[...] *= -1
    Fcc.mat[...] *= -1
    X = _herm(orc, -1, ntau, 2)
    assert orc.vie2_mat(X, F, Fcc, G0, beta, 5, CG) == 0
    G = cntr.herm_matrix(-1, ntau, 2)
    cntr.vie2_mat(G, to_gpu(cntr, F), to_gpu(cntr, Fcc), to_gpu(cntr, G0), beta, 5, cntr.CNTR_MAT_CG)
    assert mat_err(orc, G, X) <= TOL
    assert np.abs(to_host(orc, G).mat - Gx.mat).max() < 1e-11  # the exact 4x4 answer


def test_reference_matsubara_test_on_the_gpu(cntr, orc):
    # libcntr/test/matsubara.cpp:35-104 (conjugate-gradient section)
    from tests.test_oracle_cg import _reference_matsubara_case
    Gx, Sig, H, mu, beta = _reference_matsubara_case(orc)
    G = cntr.herm_matrix(-1, Gx.ntau, 1)
    cntr.dyson_mat(G, mu, func_to_gpu(cntr, H), to_gpu(cntr, Sig), beta, 5, cntr.CNTR_MAT_CG, False)
    assert cntr.distance_norm2(-1, to_gpu(cntr, Gx), G) < 1e-7
