"""The run epilogue's host parts (ccpsdv3/main.c:451-570), no GPU call:

* the eigensolver behind cpsd_rotate()/cpsd_dev_epilogue() against the
  mathematics (A V = V diag(w), V orthonormal, the spectrum of numpy's
  LAPACK eigh) and eigsrt's ordering rule (ccpsdv3/eigsrt.c:11-39) on inputs
  whose answer is fixed by hand (diagonal matrices with ties);
* the ELCIN/ELCOUT and IONOUT files against the field order of
  ccpsdv3/main.c:115-140, 218-251, 539-570.
"""
import struct

import numpy as np
import pytest

import cpsd_amd
import oracle_lib as O
from example_inputs import read_ionin


def sym(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    return a + a.T


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (5, 2), (17, 3), (64, 4), (128, 5)])
def test_eigenpairs(n, seed):
    H = sym(n, seed)
    a = H.T.reshape(-1).copy()            # column-major
    eig, V = cpsd_amd.syev_sorted(a)
    Vm = V.reshape(n, n)                  # row i = eigenvector i
    scale = np.abs(np.linalg.eigvalsh(H)).max()
    assert np.all(np.diff(eig) <= 0)                                    # eigsrt: descending
    np.testing.assert_allclose(eig, np.sort(np.linalg.eigvalsh(H))[::-1], rtol=0, atol=1e-13 * scale)
    np.testing.assert_allclose(Vm @ Vm.T, np.eye(n), rtol=0, atol=1e-13 * n)
    np.testing.assert_allclose(H @ Vm.T, Vm.T * eig, rtol=0, atol=1e-12 * scale)


def test_only_upper_triangle_is_read():
    """dsyev('V','U') references the upper triangle only."""
    n = 9
    H = sym(n, 7)
    a = H.T.reshape(-1).copy()
    junk = a.copy().reshape(n, n)          # junk[j, i] = a[i + j n]: column j
    for j in range(n):
        for i in range(j + 1, n):
            junk[j, i] = 1e3 * (i - j)     # element (i, j) below the diagonal
    e1, v1 = cpsd_amd.syev_sorted(a)
    e2, v2 = cpsd_amd.syev_sorted(junk.reshape(-1))
    np.testing.assert_array_equal(e1, e2)
    np.testing.assert_array_equal(v1, v2)


def test_eigsrt_order_with_ties():
    """diag(1, 3, 3, 2): dsyev returns ascending (1, 2, 3, 3) with e0, e3,
    e1, e2; eigsrt then takes the last of equal maxima first:
    (3, 3, 2, 1) with e2, e1, e3, e0."""
    H = np.diag([1.0, 3.0, 3.0, 2.0])
    eig, V = cpsd_amd.syev_sorted(H.T.reshape(-1).copy())
    np.testing.assert_array_equal(eig, [3.0, 3.0, 2.0, 1.0])
    E = np.eye(4)
    np.testing.assert_array_equal(V.reshape(4, 4), np.array([E[2], E[1], E[3], E[0]]))
    # the oracle's restatement of eigsrt gives the same from dsyev's order
    D, W = O.eigsrt([1.0, 2.0, 3.0, 3.0], np.array([E[0], E[3], E[1], E[2]]).reshape(-1))
    np.testing.assert_array_equal(D, eig)
    np.testing.assert_array_equal(W, V)


def test_oracle_eigsrt_is_a_descending_permutation():
    rng = np.random.default_rng(11)
    d = rng.integers(0, 5, 12).astype(float)      # many ties
    v = np.arange(144, dtype=float)
    D, W = O.eigsrt(d, v)
    assert np.all(np.diff(D) <= 0)
    assert sorted(D) == sorted(d)
    rows = W.reshape(12, 12)[:, 0] / 12           # which original column moved where
    assert sorted(rows.astype(int)) == list(range(12))
    assert np.array_equal(d[rows.astype(int)], D)


def test_syev_rejects_nonfinite():
    a = np.eye(3).reshape(-1).copy()
    a[4] = np.nan
    with pytest.raises(cpsd_amd.CpsdError):
        cpsd_amd.syev_sorted(a)


def test_elc_file_layout(tmp_path):
    """int icube, int nfft, double unit, int ispin, int ne[2], then
    (ne0+ne1) n2ft3d doubles (ccpsdv3/main.c:539-547), and read back as
    ccpsdv3/main.c:115-140 reads it."""
    nfft, ne = 8, 3
    n2 = (nfft + 2) * nfft * nfft
    c = np.random.default_rng(1).standard_normal(ne * n2)
    p = tmp_path / "ELCOUT"
    cpsd_amd.write_elc(str(p), 1, nfft, 20.0, c, ne)
    raw = p.read_bytes()
    assert len(raw) == 4 + 4 + 8 + 4 + 8 + 8 * ne * n2
    assert struct.unpack("=iidiii", raw[:28]) == (1, nfft, 20.0, 1, ne, 0)
    np.testing.assert_array_equal(np.frombuffer(raw[28:], dtype=np.float64), c)
    r = cpsd_amd.read_elc(str(p))
    assert (r["icube"], r["nfft"], r["unit"], r["ispin"], r["ne"]) == (1, nfft, 20.0, 1, [ne, 0])
    np.testing.assert_array_equal(r["c"], c)


def test_elc_read_rejects_short_file(tmp_path):
    p = tmp_path / "ELCIN"
    p.write_bytes(struct.pack("=iidiii", 1, 8, 20.0, 1, 3, 0) + b"\0" * 16)
    with pytest.raises(cpsd_amd.CpsdError):
        cpsd_amd.read_elc(str(p))


def test_ionout_format_and_ionin_round_trip(tmp_path):
    """IONOUT as ccpsdv3/main.c:551-570 writes it, read back with the IONIN
    reader of ccpsdv3/main.c:218-251."""
    p = tmp_path / "IONOUT"
    r2 = np.array([0.0, 0.0, 0.0, 1.431030, 1.108120, 0.0, -1.431030, 1.108120, -2.5])
    cpsd_amd.write_ionout(str(p), ["O00", "H00"], [2, 1], [0, 1, 1], r2)
    txt = p.read_text()
    assert txt.splitlines()[:4] == ["2", "'O00 ' 2", "'H00 ' 1", "3"]
    assert txt.splitlines()[5] == "2 1.431030 1.108120 0.000000 0.000000 0.000000 0.000000 "
    kinds, katm, r = read_ionin(str(p))
    assert kinds == [("O00", 2), ("H00", 1)]
    np.testing.assert_array_equal(katm, [0, 1, 1])
    np.testing.assert_allclose(r, r2, rtol=0, atol=5e-7)
