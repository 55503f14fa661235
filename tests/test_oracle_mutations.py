"""The oracle's pins reject plausible mistakes.

Each case copies oracle/cntr_oracle.cpp, makes one small change of the kind a restatement
could get wrong (a block product in the wrong order, a conjugate transpose dropped, a sign
of a Hermitian continuation, a transposed orbital index, an integration weight off in the
ninth digit), builds the altered copy into a temporary library, and runs the pins against
it: every pin listed for the case must fail, while the unaltered oracle passes all of them
(the pins' own test modules).
"""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

from tests import test_oracle_cg as CG
from tests import test_oracle_convolution_exact as CX
from tests import test_oracle_matrix_pins as MP
from tests import test_oracle_green_td as GT
from tests import test_oracle_physics as PH

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                   "cntr_oracle.cpp")

PINS = {
    "convolution_scalar_identity": lambda o: PH.test_convolution_single_pole_identity(o, -1),
    "convolution_size2_identity": lambda o: MP.test_convolution_size2_matrix_identity(o, -1),
    "convolution_size2_polynomial": lambda o: CX.test_convolution_size2_exact_for_polynomials(o, -1),
    "dyson_size2_embedding": lambda o: MP.test_dyson_size2_embedding_exact(o),
    "dyson_size2_order_in_h": lambda o: MP.test_dyson_size2_retarded_order_in_h(o),
    "vie2_size2_embedding": lambda o: MP.test_vie2_size2_embedding_exact(o),
    "bubble1_size2": lambda o: MP.test_bubble_size2_every_index_map(o, 1, -1, -1),
    "dyson_scalar_embedding": lambda o: PH.test_dyson_embedding_matches_reference_tolerance(o),
    "green_td_cubic_cf4": lambda o: GT.test_scalar_cubic_eps_is_exact_for_cf4(),
    "green_td_order4": lambda o: GT.test_noncommuting_2x2_converges_at_the_order(4, 16.0),
    "cg_reference_matsubara": lambda o: CG.test_reference_matsubara_test_conjugate_gradient(o),
    "cg_symmetrised_equation": lambda o: CG.test_cg_solves_the_symmetrised_equation_and_matches_fixpoint(o),
    "cg_matrix_valued": lambda o: CG.test_cg_matrix_valued_dyson_and_vie2(o),
}

# name: (text in cntr_oracle.cpp, replacement, occurrences replaced, pins that must fail)
MUTATIONS = {
    "block_product_order": ("acc += a[i * s + l] * b[l * s + j];",
                            "acc += b[i * s + l] * a[l * s + j];", 2,
                            ["convolution_size2_identity", "convolution_size2_polynomial",
                             "dyson_size2_embedding", "vie2_size2_embedding"]),
    "left_solve_without_conjugate_transpose": (
        "for (int b = 0; b < nb * nb; b++) e_herm(s, Mc.data() + b * es, M + b * es);",
        "for (int b = 0; b < nb * nb; b++) e_set(s, Mc.data() + b * es, M + b * es);", 1,
        ["dyson_size2_embedding"]),
    "retarded_continuation_sign": ("e_herm(s, bt.data(), Bcc.ret(j, m));\n                    w = -w;",
                                   "e_herm(s, bt.data(), Bcc.ret(j, m));", 1,
                                   ["convolution_scalar_identity", "convolution_size2_identity",
                                    "convolution_size2_polynomial"]),
    "lesadv_without_conjugate_transpose": (
        "e_herm(s, badv.data() + m * es, Bcc.ret(n, m));",
        "e_set(s, badv.data() + m * es, Bcc.ret(n, m));", 1,
        ["convolution_size2_identity", "convolution_size2_polynomial"]),
    "tvvt_sign": ("bt[l] *= idtau * (-(double)sig);", "bt[l] *= idtau * ((double)sig);", 1,
                  ["convolution_scalar_identity", "convolution_size2_identity",
                   "convolution_size2_polynomial"]),
    "bubble1_transposed_index": ("cplx bles21_t1t = B.les(m, tstp)[b21];",
                                 "cplx bles21_t1t = B.les(m, tstp)[b12];", 1, ["bubble1_size2"]),
    "gregory_end_weight_1e-9": ("omega[j] = om[j].d();",
                                "omega[j] = om[j].d() * (j == 2 ? 1.0 + 1e-9 : 1.0);", 1,
                                ["convolution_scalar_identity", "convolution_size2_identity",
                                 "convolution_size2_polynomial"]),
    "gregory_start_weight_1e-9": (
        "greg[n * k2 + i] = Wr[(0 * k1 + n) * k1 + i].d();",
        "greg[n * k2 + i] = Wr[(0 * k1 + n) * k1 + i].d() * (n == 3 && i == 1 ? 1.0 + 1e-9 : 1.0);", 1,
        ["convolution_scalar_identity", "convolution_size2_identity", "convolution_size2_polynomial"]),
    "matsubara_edge_correction_1e-9": ("rc[((m - 1) * k1 + j) * k1 + l] = s.d();",
                                       "rc[((m - 1) * k1 + j) * k1 + l] = s.d() * (1.0 + 1e-9);", 1,
                                       ["convolution_scalar_identity", "convolution_size2_identity",
                                        "convolution_size2_polynomial"]),
    "bdf_coefficient_1e-6": ("bd[l] = s.d();", "bd[l] = s.d() * (l == 2 ? 1.0 + 1e-6 : 1.0);", 1,
                             ["dyson_size2_order_in_h"]),
    "cf4_one_quadrature_node": (
        "func_interp(H, n1, k, t - 0.5 + std::sqrt(3.0) / 6.0, h2.data());",
        "func_interp(H, n1, k, t - 0.5 - std::sqrt(3.0) / 6.0, h2.data());", 1,
        ["green_td_cubic_cf4", "green_td_order4"]),
    "cf4_exponentials_swapped": ("e_mult(s, tmp.data(), e1.data(), e2.data());",
                                 "e_mult(s, tmp.data(), e2.data(), e1.data());", 1,
                                 ["green_td_order4"]),
    "cg_residual_without_half": ("cplx r = (C1.mat_[l] + C2.mat_[l]) * 0.5;",
                                 "cplx r = (C1.mat_[l] + C2.mat_[l]);", 1,
                                 ["cg_reference_matsubara", "cg_symmetrised_equation", "cg_matrix_valued"]),
    "cg_fcc_on_the_left": ("convolution_matsubara(C2, P, Fcc, In, beta);",
                           "convolution_matsubara(C2, Fcc, P, In, beta);", 1, ["cg_matrix_valued"]),
    "cg_sigma_g0_order": ("convolution_matsubara(SxG0, S, G0, In, beta);",
                          "convolution_matsubara(SxG0, G0, S, In, beta);", 1, ["cg_matrix_valued"]),
}


def _build(tmp, name):
    old, new, count, _ = MUTATIONS[name]
    src = open(SRC).read()
    assert src.count(old) == count, (name, src.count(old))
    d = os.path.join(tmp, name)
    os.makedirs(d, exist_ok=True)
    cpp = os.path.join(d, "cntr_oracle.cpp")
    with open(cpp, "w") as f:
        f.write(src.replace(old, new))
    inc = os.path.dirname(SRC)
    so = os.path.join(d, "liboracle_%s.so" % name.replace("-", "_"))
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-shared", "-I" + inc,
                           "-o", so, cpp])
    return so


@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("oracle_mutants"))
    with ThreadPoolExecutor(min(8, os.cpu_count() or 1)) as ex:
        libs = list(ex.map(lambda n: _build(tmp, n), MUTATIONS))
    return dict(zip(MUTATIONS, libs))


@pytest.mark.parametrize("name", list(MUTATIONS))
def test_pins_reject_mutation(orc, mutants, name):
    must_fail = MUTATIONS[name][3]
    survived = []
    with orc.using(mutants[name]):
        for pin in must_fail:
            try:
                PINS[pin](orc)
            except AssertionError:
                continue
            survived.append(pin)
    assert not survived, "%s passed pins %s" % (name, survived)
