"""The C++ drop-in (include/cntr_amd.hpp) driven by tests/cpp/dropin_driver.cpp on the GPU:
a two-orbital system with a two-orbital bath (non-commuting complex couplings) in the
Dyson and VIE2 forms against the exact 4x4 solution, a checkpoint at timestep 37 resumed
bit for bit, and herm_matrix_timestep / herm_matrix_timestep_view objects in bubbles,
distance_norm2 and timestep sums.  The Dyson-form G equals the oracle's on the same inputs
to 1e-12 per component."""
import os
import subprocess

import numpy as np
import pytest

from tests import physics
from tests.gpu_util import rel_err

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "dropin_driver")

HS = np.array([[-1.0, 0.3 + 0.2j], [0.3 - 0.2j, 0.5]])
HB = np.array([[0.8, -0.25j], [0.25j, -0.4]])
VC = np.array([[0.3, 0.1 + 0.2j], [-0.15j, 0.25]])


def test_cpp_dropin_driver(orc, tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    assert os.path.exists(BIN), "tests/cpp/bin/dropin_driver not built (run build())"
    out = tmp_path / "g.bin"
    r = subprocess.run([BIN, str(tmp_path), str(out)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)

    nt, ntau, beta, h, k, mu = 80, 300, 5.0, 0.05, 5, -0.1
    ntri, ntv = (nt + 1) * (nt + 2) // 2, (nt + 1) * (ntau + 1)
    g = np.fromfile(str(out), dtype=np.complex128).reshape(-1, 2, 2)
    assert g.shape[0] == 2 * ntri + ntv + ntau + 1
    parts = {"ret": g[:ntri], "les": g[ntri:2 * ntri], "tv": g[2 * ntri:2 * ntri + ntv],
             "mat": g[2 * ntri + ntv:]}
    gb = physics.free_matrix(orc.HostHerm(nt, ntau, 2), HB, mu, beta, h)
    Sig = orc.HostHerm(nt, ntau, 2)
    for c in ("ret", "les", "tv", "mat"):
        getattr(Sig, c)[...] = VC @ getattr(gb, c) @ VC.conj().T
    H = orc.HostFunc(nt, 2)
    H.data[...] = HS
    X = orc.HostHerm(nt, ntau, 2)
    orc.dyson_mat(X, mu, H, Sig, beta, k, 2)
    orc.dyson_start(X, mu, H, Sig, beta, h, k)
    for n in range(k + 1, nt + 1):
        orc.dyson_timestep(n, X, mu, H, Sig, beta, h, k)
    for c in ("ret", "les", "tv", "mat"):
        e = rel_err(parts[c], getattr(X, c))
        assert e <= 1e-12, (c, e)
