"""Host-side checks of the batch server that need no GPU: the footprint the C ABI reports
for a request (the admission currency), and the seeded request generator."""
import numpy as np


def herm_bytes(nt, ntau, s):
    r16 = lambda x: (x + 15) // 16 * 16
    es = s * s
    return 16 * (2 * r16((nt + 1) * (nt + 2) // 2 * es) + r16((nt + 1) * (ntau + 1) * es) +
                 r16((ntau + 1) * es))


def test_footprint_matches_layout():
    from nessi_amd import batch_server, workloads
    for r in workloads.uneven_batch(16, seed=5):
        s, nt, ntau = r["size1"], r["nt"], r["ntau"]
        g = herm_bytes(nt, ntau, s)
        r16 = lambda x: (x + 15) // 16 * 16
        aux = herm_bytes(nt, ntau, s) + 16 * r16((nt + 2) * s * s)
        rnd = lambda b: (b + 255) // 256 * 256
        assert batch_server.footprint(r) == rnd(g) + rnd(aux)
        # G and Sigma dominate: two herm_matrix of the reference's size
        ref = 16 * s * s * ((nt + 1) * (nt + 2) + (nt + 1) * (ntau + 1) + ntau + 1)
        assert 2 * ref <= batch_server.footprint(r) <= 2 * ref * 1.2 + 65536


def test_footprint_rejects_bad_sizes():
    from nessi_amd import batch_server, workloads
    r = workloads.uneven_batch(1, seed=1)[0]
    r["size1"] = 3
    r["H"] = np.zeros((r["nt"] + 2, 3, 3), np.complex128)
    r["Hbath"] = np.zeros((3, 3), np.complex128)
    assert batch_server.footprint(r) == 0


def test_uneven_batch_distribution():
    from nessi_amd import workloads
    a = workloads.uneven_batch(64, seed=9)
    b = workloads.uneven_batch(64, seed=9)
    for x, y in zip(a, b):
        assert x["size1"] == y["size1"] and np.array_equal(x["H"], y["H"])
    assert {r["size1"] for r in a} == {1, 2, 4, 8}
    for r in a:
        assert 200 <= r["nt"] <= 1500 and 2 <= r["beta"] <= 20 and r["ntau"] == 300
        H = r["H"]
        assert np.allclose(H, np.conj(np.transpose(H, (0, 2, 1))))
        assert np.array_equal(H[0], H[1])  # H(-1) = H(0) = H0 (sin 0 = 0)
        assert np.allclose(r["Hbath"], np.diag(np.diag(r["Hbath"].real)))
    t = workloads.poisson_arrivals(50.0, 4000, seed=3)
    assert np.all(np.diff(t) > 0) and abs(4000 / t[-1] - 50.0) < 5.0


def test_dyson_solve_is_the_reference_sequence(orc):
    """orc.dyson_solve composes the individually pinned oracle functions exactly"""
    from nessi_amd import workloads
    r = workloads.uneven_batch(1, seed=3, nt_lo=20, nt_hi=30, ntau=40, sizes=(2,))[0]
    G, _ = orc.dyson_solve(r)
    s, nt, ntau = r["size1"], r["nt"], r["ntau"]
    S2 = orc.HostHerm(nt, ntau, s)
    orc.green_from_H_const(S2, r["mu"], r["Hbath"], r["beta"], r["h"])
    for c in ("ret", "les", "tv", "mat"):
        getattr(S2, c)[...] *= r["V"] ** 2
    H = orc.HostFunc(nt, s)
    H.data[...] = r["H"]
    G2 = orc.HostHerm(nt, ntau, s)
    orc.dyson_mat(G2, r["mu"], H, S2, r["beta"], 5, 2, True)
    orc.dyson_start(G2, r["mu"], H, S2, r["beta"], r["h"], 5)
    for n in range(6, nt + 1):
        orc.dyson_timestep(n, G2, r["mu"], H, S2, r["beta"], r["h"], 5)
    for c in ("ret", "les", "tv", "mat"):
        assert np.array_equal(getattr(G, c), getattr(G2, c)), c
