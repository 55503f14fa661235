"""The C++ drop-in (include/cntr_amd.hpp) compiles with g++ against the C ABI and links to
libcntrx.so: the reference-shaped calls of tests/cpp/dropin_driver.cpp (herm_matrix,
herm_matrix_timestep and _view, dyson_*, vie2_*, convolution_timestep, Bubble1,
distance_norm2, text checkpoints) resolve to exported symbols.  No GPU call is made here."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "nessi_amd", "lib")


def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "cntr_amd.hpp"\n'
                   "int main() {\n"
                   "  cntr::herm_matrix<double> G(10, 20, 2, FERMION), S(G);\n"
                   "  cntr::function<double> H(10, 2);\n"
                   "  cdmatrix h = cdmatrix::Identity(2, 2) * 0.5;\n"
                   "  H.set_constant(h);\n"
                   "  G.set_timestep_zero(3); G.smul(-1, 2.0); G.set_timestep(4, S);\n"
                   "  std::complex<double> x; G.get_les(2, 5, x); G.get_vt(3, 4, x);\n"
                   "  return (int)G.density_matrix(-1).real();\n"
                   "}\n")
    r = subprocess.run(["g++", "-std=c++11", "-Wall", "-Werror", "-fsyntax-only",
                        "-I" + os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_driver_links_against_the_library(tmp_path):
    lib = os.path.join(LIBDIR, "libcntrx.so")
    if not os.path.exists(lib):
        pytest.skip("libcntrx.so not built")
    out = tmp_path / "dropin_driver"
    r = subprocess.run(["g++", "-O1", "-std=c++14", "-Wall", "-Werror",
                        "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "dropin_driver.cpp"),
                        "-L" + LIBDIR, "-lcntrx", "-Wl,-rpath," + LIBDIR, "-o", str(out)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert out.exists()
