"""Text checkpoints (herm_matrix / function print_to_file and read_from_file) on the GPU.

- The reference's own test (test/herm_matrix_readwrite.cpp:9-52): write at precision 10,
  read whole and up to nt1, distance_norm2 summed over the timesteps read < 1e-6.  (The C++
  drop-in's checkpoint and resume are covered by tests/cpp/dropin_driver.cpp.)
- The file is the reference's, byte for byte: the expected text is built here from the
  device values with the reference's print loops (cntr_herm_matrix_impl.hpp:1308-1359,
  cntr_function_impl.hpp:719-732) and Python's "%.*g", which is what an ostream of
  precision p and no floatfield writes (C++ [facet.num.put.virtuals]).
- At precision 17 a read-back is exact, so a run resumed from a checkpoint at nt1 continues
  bit for bit as the run that was never stopped.
- Files the reference would reject (short, malformed) fail with an error and leave G as it was.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def cntr():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nessi_amd import cntr as c
    return c


def _green(cntr, nt=100, ntau=50):
    """G1 of the reference's read/write test: green_from_H of a 2x2 Hamiltonian"""
    h0 = np.array([[-0.4, 0.1j], [-0.1j, 0.6]], np.complex128)
    G = cntr.herm_matrix(nt, ntau, 2, cntr.FERMION)
    cntr.green_from_H(G, 0.0, h0, 10.0, 0.01)
    return G


def _fmt(p, z):
    return " %s %s" % ("%.*g" % (p, z.real), "%.*g" % (p, z.imag))


def _expected_herm(G, p):
    """the reference's print_to_file loops over the device values"""
    a = {c: getattr(G, c).cpu().numpy().reshape(-1, G.size1 * G.size1)
         for c in ("ret", "les", "tv", "mat")}
    nt, ntau = G.nt, G.ntau
    out = ["# %d %d %d  %d\n" % (nt, ntau, G.size1, G.sig)]
    for j in range(ntau + 1):
        out.append("mat: %d%s\n" % (j, "".join(_fmt(p, z) for z in a["mat"][j])))
    out.append("\n")
    if nt >= 0:
        for i in range(nt + 1):
            for j in range(i + 1):
                out.append("ret: %d %d%s\n" % (i, j, "".join(_fmt(p, z) for z in a["ret"][i * (i + 1) // 2 + j])))
            out.append("\n")
        out.append("\n")
        for i in range(nt + 1):
            for j in range(ntau + 1):
                out.append("tv: %d %d%s\n" % (i, j, "".join(_fmt(p, z) for z in a["tv"][i * (ntau + 1) + j])))
            out.append("\n")
        out.append("\n")
        for j in range(nt + 1):
            for i in range(j + 1):
                out.append("les: %d %d%s\n" % (i, j, "".join(_fmt(p, z) for z in a["les"][j * (j + 1) // 2 + i])))
            out.append("\n")
        out.append("\n")
    return "".join(out)


def test_reference_readwrite_test(cntr, tmp_path):
    f = str(tmp_path / "herm_matrix_plain.dat")
    G1 = _green(cntr)
    G1.print_to_file(f, 10)
    G2 = cntr.herm_matrix(100, 50, 2, cntr.FERMION)
    G2.read_from_file(f)
    err = sum(cntr.distance_norm2(t, G1, G2) for t in range(-1, 101))
    assert err < 1e-6, err
    G3 = cntr.herm_matrix(100, 50, 2, cntr.FERMION)
    G3.read_from_file(50, f)
    err = sum(cntr.distance_norm2(t, G1, G3) for t in range(-1, 51))
    assert err < 1e-6, err
    # timesteps after nt1 are left as they were (zero)
    assert float(G3.ret[51 * 52 // 2:].abs().max()) == 0.0
    assert float(G3.tv[51 * 51:].abs().max()) == 0.0
    assert float(G3.les[51 * 52 // 2:].abs().max()) == 0.0


@pytest.mark.parametrize("p", [10, 16, 17])
def test_file_is_the_reference_format(cntr, tmp_path, p):
    G = _green(cntr, nt=12, ntau=9)
    G.sig = G._desc.sig = cntr.FERMION
    f = tmp_path / "g.dat"
    G.print_to_file(str(f), p)
    assert f.read_text() == _expected_herm(G, p)
    Gm = cntr.herm_matrix(-1, 9, 2, cntr.BOSON)  # Matsubara only: header and mat block
    Gm.mat.copy_(G.mat)
    Gm.print_to_file(str(f), p)
    assert f.read_text() == _expected_herm(Gm, p)


def test_function_file_format_and_round_trip(cntr, tmp_path):
    import torch
    rng = np.random.default_rng(5)
    F = cntr.function(7, 2)
    F.data.copy_(torch.from_numpy(rng.standard_normal((9, 2, 2)) + 1j * rng.standard_normal((9, 2, 2))))
    f = tmp_path / "f.dat"
    F.print_to_file(str(f), 12)
    vals = F.data.cpu().numpy().reshape(9, 4)
    exp = "# 7 2\n" + "".join(
        "".join("%s %s " % ("%.*g" % (12, z.real), "%.*g" % (12, z.imag)) for z in row) + "\n"
        for row in vals)
    assert f.read_text() == exp
    F.print_to_file(str(f), 17)
    F2 = cntr.function(3, 1)  # shorter and of another size1: the reader reallocates
    F2.read_from_file(str(f))
    assert (F2.nt, F2.size1) == (7, 2)
    assert torch.equal(F2.data, F.data)
    F3 = cntr.function(7, 2)
    F3.read_from_file(4, str(f))
    assert torch.equal(F3.data[:6], F.data[:6])
    assert float(F3.data[6:].abs().max()) == 0.0


def test_precision17_is_exact_and_resizes(cntr, tmp_path):
    import torch
    f = str(tmp_path / "g.dat")
    G = _green(cntr, nt=30, ntau=20)
    G.print_to_file(f, 17)
    G2 = cntr.herm_matrix(30, 20, 2, cntr.FERMION)
    G2.read_from_file(f)
    assert torch.equal(G2._storage, G._storage)
    G3 = cntr.herm_matrix(5, 7, 1, cntr.BOSON)  # another shape and sig: reallocated
    G3.read_from_file(f)
    assert (G3.nt, G3.ntau, G3.size1, G3.sig) == (30, 20, 2, cntr.FERMION)
    assert torch.equal(G3._storage, G._storage)


@pytest.mark.parametrize("size1", [1, 4, 8])
def test_every_size_round_trips_and_matches_the_format(cntr, tmp_path, size1):
    """size1 1, 4, 8 (blocks written row-major per element): the bytes against the
    reference's loops at 16 digits, and an exact read-back at 17"""
    import torch
    rng = np.random.default_rng(size1)
    X = rng.standard_normal((size1, size1)) + 1j * rng.standard_normal((size1, size1))
    G = cntr.herm_matrix(9, 11, size1, cntr.BOSON if size1 == 4 else cntr.FERMION)
    cntr.green_from_H(G, 0.2, (X + X.conj().T) / 2, 4.0, 0.05)
    f = tmp_path / "g.dat"
    G.print_to_file(str(f), 16)
    assert f.read_text() == _expected_herm(G, 16)
    G.print_to_file(str(f), 17)
    G2 = cntr.herm_matrix(9, 11, size1, cntr.FERMION)
    G2.read_from_file(str(f))
    assert G2.sig == G.sig
    assert torch.equal(G2._storage, G._storage)


def test_resume_from_checkpoint_is_bit_exact(cntr, tmp_path):
    """dyson(...) to nt; the checkpoint read up to nt1 and continued with dyson_timestep
    nt1+1..nt gives the same G, bit for bit"""
    import torch
    nt, ntau, k, beta, h, mu = 40, 60, 5, 5.0, 0.02, 0.1
    Hb = np.array([[0.5, 0.2], [0.2, -0.7]], np.complex128)
    H = cntr.function(nt, 2)
    H.set_constant(np.array([[-0.3, 0.1j], [-0.1j, 0.2]]))
    Sig = cntr.herm_matrix(nt, ntau, 2, cntr.FERMION)
    cntr.green_from_H(Sig, mu, Hb, beta, h)
    Sig._storage.mul_(0.16)
    G = cntr.herm_matrix(nt, ntau, 2, cntr.FERMION)
    cntr.dyson(G, mu, H, Sig, beta, h, k)
    f = str(tmp_path / "ckpt.dat")
    G.print_to_file(f, 17)
    G2 = cntr.herm_matrix(nt, ntau, 2, cntr.FERMION)
    G2.read_from_file(20, f)
    for n in range(21, nt + 1):
        cntr.dyson_timestep(n, G2, mu, H, Sig, beta, h, k)
    assert torch.equal(G2._storage, G._storage)


def test_bad_files_fail_and_leave_g_unchanged(cntr, tmp_path):
    import torch
    from nessi_amd._lib import CntrxError
    G = _green(cntr, nt=8, ntau=6)
    f = tmp_path / "g.dat"
    G.print_to_file(str(f), 17)
    text = f.read_text()
    G2 = _green(cntr, nt=8, ntau=6)
    G2._storage.mul_(2.0)
    before = G2._storage.clone()
    short = tmp_path / "short.dat"
    short.write_text(text[: len(text) * 2 // 3])
    with pytest.raises(CntrxError, match="error at"):
        G2.read_from_file(str(short))
    bad = tmp_path / "bad.dat"
    bad.write_text(text.replace("tv: 3 2 ", "tv: 3 2 x", 1))
    with pytest.raises(CntrxError, match=r"error at tv \(3,2\)"):
        G2.read_from_file(str(bad))
    with pytest.raises(CntrxError, match="nt1 > n"):
        cntr.herm_matrix(10, 6, 2).read_from_file(9, str(f))
    with pytest.raises(CntrxError, match="m /= ntau_"):
        cntr.herm_matrix(8, 7, 2).read_from_file(4, str(f))
    with pytest.raises(CntrxError, match="cannot open"):
        G2.read_from_file(str(tmp_path / "missing.dat"))
    assert torch.equal(G2._storage, before)
