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


def _oracle_solve(orc, r, nt_stop):
    from oracle import oracle
    X, _ = oracle.dyson_solve(r, nt_stop=nt_stop)
    return X


def test_uneven_and_serving_shapes_alone_equal_mix(bs, orc):
    """Requests of the uneven-batch and serving workloads at full length (one of each size1
    from the seed-2024 batch, nt 200..1500, ntau 300, and two of the serving distribution
    with nt up to 2000), half of them joining 25 ticks late: each result is the same bit
    for bit as the request run alone; oracle parity over timesteps 0..40 for all, and
    over the whole propagation where the oracle is quick (size1 <= 2, nt <= 400)."""
    batch_server, workloads = bs
    pool = workloads.uneven_batch(64, seed=2024)
    rs = [next(r for r in pool if r["size1"] == s) for s in (1, 2, 4, 8)]
    rs += workloads.uneven_batch(2, seed=101, nt_lo=1600, nt_hi=2000)
    print("sizes", [r["size1"] for r in rs], "nt", [r["nt"] for r in rs], flush=True)
    S = batch_server.BatchServer(64 << 30)
    ids = [S.submit(r) for r in rs[0::2]]
    for _ in range(25):
        S.step()
    ids_late = [S.submit(r) for r in rs[1::2]]
    S.drain()
    order = [None] * len(rs)
    order[0::2], order[1::2] = ids, ids_late
    mixed = [S.result(i).to_numpy() for i in order]
    for r, m in zip(rs, mixed):
        A = batch_server.BatchServer(64 << 30)
        a = A.submit(r)
        A.drain()
        alone = A.result(a).to_numpy()
        A.close()
        for c in ("ret", "les", "tv", "mat"):
            assert np.array_equal(alone[c], m[c]), (r["size1"], r["nt"], c)
    for r, m in zip(rs, mixed):
        full = r["size1"] <= 2 and r["nt"] <= 400
        nstop = r["nt"] if full else 40
        X = _oracle_solve(orc, r, nstop)
        o = (nstop + 1) * (nstop + 2) // 2
        ntv = (nstop + 1) * (r["ntau"] + 1)
        for c, nn in (("ret", o), ("les", o), ("tv", ntv), ("mat", r["ntau"] + 1)):
            ref = getattr(X, c)[:nn]
            e = np.abs(m[c][:nn] - ref).max() / np.abs(ref).max()
            assert e <= 1e-12, (r["size1"], r["nt"], c, e)
    S.close()
