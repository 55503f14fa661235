"""The reference's Kadanoff-Baym test (test/dyson.cpp:23-119) compiled against the C++
drop-in (include/cntr_amd.hpp) and run on the GPU: both forms meet the reference's own
tolerances against the exact 2x2 solution, and the Dyson-form G equals the oracle's on
the same inputs to 1e-12 per component."""
import os
import subprocess

import numpy as np
import pytest

from tests.gpu_util import rel_err

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "dyson_embedding")


def test_reference_dyson_test_on_the_cpp_dropin(orc, tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    assert os.path.exists(BIN), "tests/cpp/bin/dyson_embedding not built (run build())"
    out = tmp_path / "g.bin"
    r = subprocess.run([BIN, str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)

    nt, ntau, beta, h, k, mu = 200, 400, 10.0, 0.05, 5, -0.1
    eps1, eps2, lam = -1.0, 1.0, 0.2
    ntri, ntv = (nt + 1) * (nt + 2) // 2, (nt + 1) * (ntau + 1)
    g = np.fromfile(str(out), dtype=np.complex128)
    assert g.size == 2 * ntri + ntv + ntau + 1
    parts = {"ret": g[:ntri], "les": g[ntri:2 * ntri], "tv": g[2 * ntri:2 * ntri + ntv],
             "mat": g[2 * ntri + ntv:]}

    H = orc.HostFunc(nt, 1)
    H.data[...] = eps1
    Sig = orc.HostHerm(nt, ntau, 1)
    orc.green_from_H_const(Sig, mu, [[eps2]], beta, h)
    for c in ("ret", "les", "tv", "mat"):
        getattr(Sig, c)[...] *= lam * lam
    X = orc.HostHerm(nt, ntau, 1)
    orc.dyson_mat(X, mu, H, Sig, beta, k, 2)
    orc.dyson_start(X, mu, H, Sig, beta, h, k)
    for n in range(k + 1, nt + 1):
        orc.dyson_timestep(n, X, mu, H, Sig, beta, h, k)
    for c in ("ret", "les", "tv", "mat"):
        e = rel_err(parts[c], getattr(X, c).ravel())
        assert e <= 1e-12, (c, e)
