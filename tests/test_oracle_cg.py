"""Pins of the oracle's conjugate-gradient Matsubara solver (CNTR_MAT_CG,
libcntr/cntr/cntr_vie2_impl.hpp:485-602, cntr_dyson_impl.hpp:522-540):
  * the reference's own test (libcntr/test/matsubara.cpp:34-104): a level embedded in a
    2-level system, dyson_mat by CG against the exact G^M within its tolerance 1e-7;
  * the CG result solves the symmetrised discrete equation
    G + (F*G + G*Fcc)/2 = G0 to near round-off and agrees with the fixpoint solution;
  * matrix-valued: a 2-level system with a 2-level bath and non-commuting complex couplings,
    dyson_mat and vie2_mat by CG against G_ss of the exact 4x4 free GF.
"""
import numpy as np

from tests import physics
from tests.test_oracle_matrix_pins import HS, _embedded, _herm

CG = 1


def _reference_matsubara_case(orc, Ntau=400, beta=20.0):
    # libcntr/test/matsubara.cpp:35-76
    eps1, eps2, lam, mu = -1.0, 1.0, 0.2, 0.0
    h2 = np.array([[eps1, 1j * lam], [-1j * lam, eps2]])
    G2 = orc.HostHerm(-1, Ntau, 2)
    orc.green_from_H_const(G2, mu, h2, beta, 1.0)
    Gx = orc.HostHerm(-1, Ntau, 1)
    Gx.mat[:, 0, 0] = G2.mat[:, 0, 0]
    Sig = orc.HostHerm(-1, Ntau, 1)
    orc.green_from_H_const(Sig, mu, [[eps2]], beta, 1.0)
    Sig.mat[...] *= lam * lam
    H = orc.HostFunc(-1, 1)
    H[-1] = eps1
    return Gx, Sig, H, mu, beta


def test_reference_matsubara_test_conjugate_gradient(orc):
    Gx, Sig, H, mu, beta = _reference_matsubara_case(orc)
    G = orc.HostHerm(-1, Gx.ntau, 1)
    assert orc.dyson_mat(G, mu, H, Sig, beta, 5, CG, False) == 0
    err = orc.distance_norm2(-1, Gx, G)
    assert err < 1e-7  # libcntr/test/matsubara.cpp:103
    # measured 1.06e-9: the discretisation error of the Gregory convolution at ntau 400
    assert err < 3e-9, err


def test_cg_solves_the_symmetrised_equation_and_matches_fixpoint(orc):
    Gx, Sig, H, mu, beta = _reference_matsubara_case(orc, Ntau=300, beta=10.0)
    Ntau = Gx.ntau
    G = orc.HostHerm(-1, Ntau, 1)
    orc.dyson_mat(G, mu, H, Sig, beta, 5, CG, False)
    G0 = orc.HostHerm(-1, Ntau, 1)
    orc.green_from_H_const(G0, mu, [[-1.0]], beta, 1.0)
    F, Fcc, C1, C2 = (orc.HostHerm(-1, Ntau, 1) for _ in range(4))
    orc.convolution_matsubara(F, G0, Sig, beta, 5)
    orc.convolution_matsubara(Fcc, Sig, G0, beta, 5)
    F.mat[...] *= -1
    Fcc.mat[...] *= -1
    orc.convolution_matsubara(C1, F, G, beta, 5)
    orc.convolution_matsubara(C2, G, Fcc, beta, 5)
    res = G.mat + 0.5 * (C1.mat + C2.mat) - G0.mat
    assert np.abs(res).max() < 1e-13, np.abs(res).max()
    Gf = orc.HostHerm(-1, Ntau, 1)
    orc.dyson_mat(Gf, mu, H, Sig, beta, 5, 2, False)
    assert np.abs(G.mat - Gf.mat).max() < 1e-11


def test_cg_matrix_valued_dyson_and_vie2(orc):
    ntau, beta = 400, 5.0
    Gx, Sig, Hf, mu = _embedded(orc, 4, ntau, 0.05, beta)
    G = _herm(orc, -1, ntau, 2)
    S = _herm(orc, -1, ntau, 2)
    S.mat[...] = Sig.mat
    H = orc.HostFunc(-1, 2)
    H[-1] = HS
    assert orc.dyson_mat(G, mu, H, S, beta, 5, CG, False) == 0
    # measured about 2e-13, as for the fixpoint solution of the same system
    assert np.abs(G.mat - Gx.mat).max() < 1e-11, np.abs(G.mat - Gx.mat).max()
    # vie2_mat by CG: [1 - G0*Sigma] G = G0 with the explicit F and Fcc
    G0 = physics.free_matrix(_herm(orc, -1, ntau, 2), HS, mu, beta, 0.05)
    F, Fcc = _herm(orc, -1, ntau, 2), _herm(orc, -1, ntau, 2)
    orc.convolution_matsubara(F, G0, S, beta, 5)
    orc.convolution_matsubara(Fcc, S, G0, beta, 5)
    F.mat[...] *= -1
    Fcc.mat[...] *= -1
    Gv = _herm(orc, -1, ntau, 2)
    assert orc.vie2_mat(Gv, F, Fcc, G0, beta, 5, CG) == 0
    assert np.abs(Gv.mat - Gx.mat).max() < 1e-11, np.abs(Gv.mat - Gx.mat).max()
