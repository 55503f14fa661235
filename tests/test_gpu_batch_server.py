"""Continuous batching over the per-solve entry points (nessi_amd/batch_server.py): every
request's result is the same alone and in a mix of sizes and timesteps, bit for bit, and
matches the oracle's dyson(...) on the same request (<= 1e-12 per component)."""
import numpy as np
import pytest

from tests.gpu_util import component_errors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bs():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nessi_amd import _lib, batch_server, workloads
    _lib.check(_lib.load().cntrx_init(0))
    return batch_server, workloads


def _requests(workloads):
    # sizes 1, 8, 1, 4, 2, 2 and nt 17..39 (seeded)
    return workloads.uneven_batch(6, seed=11, nt_lo=14, nt_hi=40, ntau=20, h=0.02)


def test_alone_equals_in_the_mix(bs):
    batch_server, workloads = bs
    rs = _requests(workloads)
    print("sizes", [r["size1"] for r in rs], "nt", [r["nt"] for r in rs], flush=True)
    S = batch_server.BatchServer(8 << 30)
    ids = [S.submit(r) for r in rs[:3]]
    for _ in range(7):  # the later requests join at other timesteps
        S.step()
    ids += [S.submit(r) for r in rs[3:]]
    S.drain()
    done = dict(S.poll())
    assert sorted(done) == sorted(ids)
    mixed = [S.result(i).to_numpy() for i in ids]
    for r, m in zip(rs, mixed):
        A = batch_server.BatchServer(8 << 30)
        a = A.submit(r)
        A.drain()
        alone = A.result(a).to_numpy()
        for c in ("ret", "les", "tv", "mat"):
            assert np.array_equal(alone[c], m[c]), (r["size1"], r["nt"], c)


def test_server_matches_oracle(bs, orc):
    batch_server, workloads = bs
    rs = _requests(workloads)
    S = batch_server.BatchServer(8 << 30)
    ids = [S.submit(r) for r in rs]
    S.drain()
    for r, rid in zip(rs, ids):
        s, nt, ntau, k = r["size1"], r["nt"], r["ntau"], r["SolveOrder"]
        H = orc.HostFunc(nt, s)
        H.data[...] = r["H"]
        Sig = orc.HostHerm(nt, ntau, s)
        orc.green_from_H_const(Sig, r["mu"], r["Hbath"], r["beta"], r["h"])
        for c in ("ret", "les", "tv", "mat"):
            getattr(Sig, c)[...] *= r["V"] ** 2
        X = orc.HostHerm(nt, ntau, s)
        orc.dyson_mat(X, r["mu"], H, Sig, r["beta"], k, 2)
        orc.dyson_start(X, r["mu"], H, Sig, r["beta"], r["h"], k)
        for n in range(k + 1, nt + 1):
            orc.dyson_timestep(n, X, r["mu"], H, Sig, r["beta"], r["h"], k)
        errs = component_errors(orc, S.result(rid), X)
        assert all(e <= 1e-12 for e in errs.values()), (s, nt, errs)
