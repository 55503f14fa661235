"""k-batched lattice (the 1D-chain second-order workload) on the GPU against the CPU
oracle's OpenMP lattice run, plus batched == plain bit for bit."""
import numpy as np
import pytest

from tests.gpu_util import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lat_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nessi_amd import lattice
    return lattice


def run_gpu(lat_mod, epsk, U, beta, ntau, nt, h, k, nstop, batched=True):
    L = lat_mod.Lattice(epsk, U, beta, ntau, nt, h, k)
    L.equilibrium()
    iters = L.bootstrap(1e-12, 50)
    for t in range(k + 1, nstop + 1):
        it, err = L.step(t, batched, 1e-12, 50)
        iters += it
    return L, iters


def test_lattice_matches_oracle(lat_mod, orc):
    nk, U, beta, ntau, nt, h, k = 8, 2.0, 5.0, 40, 30, 0.025, 5
    epsk = lat_mod.chain_1d(nk)
    L, iters = run_gpu(lat_mod, epsk, U, beta, ntau, nt, h, k, nt)
    dens, it_cpu, _, gl = orc.lattice_run(epsk, U, beta, ntau, nt, h, k, nt, 1e-12, 50, 4, True)
    g = L.fetch(-1)
    for c in ("ret", "les", "tv"):
        e = rel_err(g[c], getattr(gl, c))
        assert e <= 1e-12, (c, e)
    dgpu = np.array([(-1j * g["les"][t * (t + 1) // 2 + t, 0, 0]).real for t in range(nt + 1)])
    assert np.abs(dgpu - dens).max() <= 1e-12
    L.close()


def test_lattice_batched_equals_plain(lat_mod):
    nk, U, beta, ntau, nt, h, k = 6, 2.0, 5.0, 30, 16, 0.025, 5
    epsk = lat_mod.chain_1d(nk)
    A, ia = run_gpu(lat_mod, epsk, U, beta, ntau, nt, h, k, nt, True)
    B, ib = run_gpu(lat_mod, epsk, U, beta, ntau, nt, h, k, nt, False)
    assert ia == ib
    for q in list(range(nk)) + [-1]:
        a, b = A.fetch(q), B.fetch(q)
        for c in ("ret", "les", "tv", "mat"):
            assert np.array_equal(a[c], b[c]), (q, c)
    A.close()
    B.close()
