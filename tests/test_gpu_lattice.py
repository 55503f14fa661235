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


@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("sampled", [False, True])
def test_lattice_run_equals_steps(lat_mod, sampled, graphs, monkeypatch):
    """The device-driven run over many timesteps gives the per-timestep loop's results bit
    for bit, with the same corrector iterations per timestep, in two calls back to back;
    with the sampled profile its chunks are split around the timed kernel.  The run's
    chunks are launched kernel by kernel, or with CNTRX_LATTICE_GRAPHS=1 at create as
    captured graphs."""
    nk, U, beta, ntau, nt, h, k = 12, 2.0, 5.0, 30, 48, 0.025, 5
    epsk = lat_mod.chain_1d(nk)
    A = lat_mod.Lattice(epsk, U, beta, ntau, nt, h, k)
    if graphs:
        monkeypatch.setenv("CNTRX_LATTICE_GRAPHS", "1")
    B = lat_mod.Lattice(epsk, U, beta, ntau, nt, h, k)
    monkeypatch.delenv("CNTRX_LATTICE_GRAPHS", raising=False)
    for L in (A, B):
        L.equilibrium()
        L.bootstrap(1e-12, 50)
    its = [A.step(t, True, 1e-12, 50)[0] for t in range(k + 1, nt + 1)]
    if sampled:
        B.profile(B.PROFILE_SAMPLED)
    mid = 20
    r1, e1 = B.run(k + 1, mid, 1e-12, 50)
    r2, e2 = B.run(mid + 1, nt, 1e-12, 50)
    assert list(r1) + list(r2) == its
    assert all(e < 1e-12 for e in list(e1) + list(e2))
    for q in (0, nk // 2, -1):
        a, b = A.fetch(q), B.fetch(q)
        for c in ("ret", "les", "tv", "mat"):
            assert np.array_equal(a[c], b[c]), (q, c)
    if sampled:
        kq, kt, kms = B.kernel_samples()
        assert len(kq) >= 8 and set(kq.tolist()) == {0, 1, 2, 3}
        assert kt.min() >= k + 1 and kt.max() <= nt and (kms > 0).all()
    A.close()
    B.close()


def test_lattice_run_hits_maxiter(lat_mod):
    """With maxiter below what convergence needs, every timestep stops at maxiter, as the
    per-timestep loop does (the stopping rule of the driver loop)."""
    nk, U, beta, ntau, nt, h, k = 6, 2.0, 5.0, 20, 14, 0.025, 5
    epsk = lat_mod.chain_1d(nk)
    A = lat_mod.Lattice(epsk, U, beta, ntau, nt, h, k)
    B = lat_mod.Lattice(epsk, U, beta, ntau, nt, h, k)
    for L in (A, B):
        L.equilibrium()
        L.bootstrap(1e-12, 50)
    ia = [A.step(t, True, 1e-12, 2) for t in range(k + 1, nt + 1)]
    ib, eb = B.run(k + 1, nt, 1e-12, 2)
    assert [x[0] for x in ia] == list(ib) and max(ib) == 2
    # the distance is summed with atomics across workgroups: equal up to rounding
    np.testing.assert_allclose([x[1] for x in ia], eb, rtol=1e-12)
    a, b = A.fetch(-1), B.fetch(-1)
    for c in ("ret", "les", "tv"):
        assert np.array_equal(a[c], b[c]), c
    A.close()
    B.close()


def test_lattice_resumes_from_checkpoint_bit_exact(lat_mod, tmp_path):
    """A run stopped at timestep 20, written as text checkpoints (17 digits) and read into a
    fresh lattice continues to timestep 36 exactly as the run that never stopped."""
    nk, U, beta, ntau, nt, h, k = 8, 2.0, 5.0, 40, 36, 0.025, 5
    epsk = lat_mod.chain_1d(nk)
    A = lat_mod.Lattice(epsk, U, beta, ntau, nt, h, k)
    A.equilibrium()
    A.bootstrap(1e-12, 50)
    A.run(k + 1, 20, 1e-12, 50)
    A.write_checkpoint(str(tmp_path))
    A.run(21, nt, 1e-12, 50)
    B = lat_mod.Lattice(epsk, U, beta, ntau, nt, h, k)
    B.read_checkpoint(str(tmp_path), 20)
    B.run(21, nt, 1e-12, 50)
    for q in list(range(nk)) + [-1, -2]:
        a, b = A.fetch(q), B.fetch(q)
        for c in ("ret", "les", "tv", "mat"):
            assert np.array_equal(a[c], b[c]), (q, c)
    A.close()
    B.close()


def test_lattice_benchmarked_workload_matches_oracle(lat_mod, orc):
    """The benchmarked workload itself (1D chain, Nk 128, beta 10, ntau 400, nt 400, h 0.025,
    U 0 -> 2, SolveOrder 5, self-consistent to 1e-12) through the device-driven run over all
    timesteps 6..400, against the oracle's OpenMP lattice over the same 400 timesteps: G_loc
    (ret, les, tv; its Matsubara part is not formed) and two G_k (every component)."""
    import os
    nk, U, beta, ntau, nt, h, k = 128, 2.0, 10.0, 400, 400, 0.025, 5
    epsk = lat_mod.chain_1d(nk)
    L = lat_mod.Lattice(epsk, U, beta, ntau, nt, h, k)
    L.equilibrium()
    L.bootstrap(1e-12, 50)
    its, _ = L.run(k + 1, nt, 1e-12, 50)
    kidx = (0, 37)
    thr = max(1, min(16, len(os.sched_getaffinity(0))))
    print("oracle lattice, %d threads" % thr, flush=True)
    _, it_cpu, _, gl, gks = orc.lattice_run(epsk, U, beta, ntau, nt, h, k, nt, 1e-12, 50, thr, True, kidx=kidx)
    g = L.fetch(-1)
    for c in ("ret", "les", "tv"):
        e = rel_err(g[c], getattr(gl, c))
        assert e <= 1e-12, ("G_loc", c, e)
    for q, X in zip(kidx, gks):
        gq = L.fetch(q)
        for c in ("ret", "les", "tv", "mat"):
            e = rel_err(gq[c], getattr(X, c))
            assert e <= 1e-12, ("G_k", q, c, e)
    L.close()
