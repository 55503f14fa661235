"""The autotuner of the batched calls (cntrx_autotune): with it on, batched timesteps of
size1 2 and 4 solves choose the recurrence panel size, and a size1 8 batch with a shared
Sigma chooses the split-K target of its GEMMs, by timing the candidates on the live jobs;
the decisions are recorded, and every result still matches the oracle to 1e-12."""
import numpy as np
import pytest

from tests.gpu_util import component_errors, func_to_gpu, to_gpu

pytestmark = pytest.mark.gpu


def _rand_herm(rng, s, scale=0.5):
    X = rng.normal(scale=scale, size=(s, s)) + 1j * rng.normal(scale=scale, size=(s, s))
    return 0.5 * (X + X.conj().T)


def test_autotuner_decides_and_results_match_oracle(orc):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nessi_amd import _lib, batch_server, cntr
    _lib.check(_lib.load().cntrx_init(0))
    batch_server.autotune(True)
    try:
        rng = np.random.default_rng(4242)
        beta, h, k, nt, ntau = 3.0, 0.04, 5, 70, 24
        for s, shared in ((2, False), (4, False), (8, True)):
            Sig = orc.HostHerm(nt, ntau, s)
            orc.green_from_H_const(Sig, 0.0, _rand_herm(rng, s), beta, h)
            for c in ("ret", "les", "tv", "mat"):
                getattr(Sig, c)[...] *= 0.1
            Hs, Xs = [], []
            for q in range(3):
                H = orc.HostFunc(nt, s)
                H0, H1 = _rand_herm(rng, s), _rand_herm(rng, s, 0.2)
                H[-1] = H0
                for t in range(nt + 1):
                    H[t] = H0 + np.sin((0.5 + 0.2 * q) * t * h) * H1
                X = orc.HostHerm(nt, ntau, s)
                orc.dyson_mat(X, 0.0, H, Sig, beta, k, 2)
                orc.dyson_start(X, 0.0, H, Sig, beta, h, k)
                for t in range(k + 1, nt + 1):
                    orc.dyson_timestep(t, X, 0.0, H, Sig, beta, h, k)
                Hs.append(H)
                Xs.append(X)
            Sg = to_gpu(cntr, Sig)
            Sgs = [Sg] * 3 if shared else [to_gpu(cntr, Sig) for _ in range(3)]
            Hg = [func_to_gpu(cntr, H) for H in Hs]
            Gs = [to_gpu(cntr, X) for X in Xs]
            cntr.dyson_timestep_batched(nt, Gs, 0.0, Hg, Sgs, beta, h, k)
            for G, X in zip(Gs, Xs):
                errs = component_errors(orc, G, X, tsteps=[nt])
                assert max(errs.values()) <= 1e-12, (s, errs)
        dec = batch_server.autotune_decisions()
        print(dec)
        got = {(d["size1"], d["tuned"]) for d in dec}
        assert (2, "recurrence panel size") in got and (4, "recurrence panel size") in got
        assert (8, "GEMM split-K workgroup target") in got
        for d in dec:
            assert all(m > 0 for m in d["ms"]) and 0 <= d["choice"] < len(d["ms"])
    finally:
        batch_server.autotune(False)
