"""GPU tests of the CSR files: MatrixMarket (SpMat.matrixmarket_write /
_write_symmetric / _read, pyqrusty/src/lib.rs:216-259) and qrusty's raw
binary format (qrusty/src/rawio.rs:128-179), plain and gzip.

Bar: the written bytes equal the oracle's restatement of each format
exactly; files read back into bit-identical device arrays; files written by
scipy (every symmetry mode) read into scipy's own matrix bit for bit.  Also
the reference's own tests (pyqrusty/tests/test_it.py:83-101, 149-198),
rewritten against this package with Python's gzip in place of the gzip
command.
"""
import gzip
import os
import shutil

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

I = np.array([[1.0, 0.0], [0.0, 1.0]], dtype=complex)
Y = np.array([[0.0, -1.0j], [1.0j, 0.0]], dtype=complex)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import qrusty_amd  # noqa: F401
    yield
    torch.cuda.synchronize()


def _op(labels, coeffs):
    import qrusty_amd as qa
    return qa.SparsePauliOp([qa.Pauli(l) for l in labels], coeffs)


def _host(m):
    _, data, indices, indptr = m.__copy__().export()
    return indptr, indices, data


def _same(a, b):
    for x, y in zip(_host(a), _host(b)):
        assert np.array_equal(np.asarray(x).view(np.uint64), np.asarray(y).view(np.uint64))


def _gunzip(path):
    with gzip.open(path, "rb") as f, open(path[:-3], "wb") as g:
        shutil.copyfileobj(f, g)
    os.remove(path)


def _gzip(path):
    with open(path, "rb") as f, gzip.open(path + ".gz", "wb") as g:
        shutil.copyfileobj(f, g)
    os.remove(path)


def test_reference_I_plus_Y(tmp_path):
    """test_it.py:83-101."""
    import qrusty_amd as qa
    from scipy.io import mmread, mmwrite
    spmat = _op(["I", "Y"], [1.0 + 0.0j, 2.0 + 0.0j]).to_matrix()
    target, target2 = str(tmp_path / "I_plus_Y.mtx"), str(tmp_path / "I_plus_Y-2.mtx")
    spmat.matrixmarket_write(target)
    spmat2 = qa.csr_matrix(spmat)
    mmwrite(target2, spmat2)
    spmat3 = mmread(target)
    assert np.array_equal(I + 2.0 * Y, spmat2.todense())
    assert np.array_equal(I + 2.0 * Y, spmat3.todense())


def test_reference_write(tmp_path):
    """test_it.py:149-198: write / read round trips through scipy, general
    and symmetric modes, plain and gzip."""
    import copy

    import qrusty_amd as qa
    from scipy.io import mmread, mmwrite

    def write_roundtrip(path, m):
        m = copy.copy(m)
        m.matrixmarket_write(str(path))
        m2 = mmread(str(path))
        assert np.array_equal(qa.csr_matrix(m).todense(), m2.todense())

    def write_roundtrip_gzip(path, m):
        gzpath = str(path) + ".gz"
        m = copy.copy(m)
        m.matrixmarket_write(gzpath)
        _gunzip(gzpath)
        m2 = mmread(str(path))
        assert np.array_equal(qa.csr_matrix(m).todense(), m2.todense())

    def write_roundtrip2(path, spmat, symmetry=None):
        mmwrite(str(path), spmat, symmetry=symmetry)
        spmat2 = mmread(str(path))
        assert np.array_equal(spmat2.todense(), spmat.todense())
        m = qa.SpMat.matrixmarket_read(str(path))
        assert np.array_equal(qa.csr_matrix(m).todense(), spmat.todense())

    def write_roundtrip2_gzip(path, spmat, symmetry=None):
        mmwrite(str(path), spmat, symmetry=symmetry)
        spmat2 = mmread(str(path))
        assert np.array_equal(spmat2.todense(), spmat.todense())
        _gzip(str(path))
        m = qa.SpMat.matrixmarket_read(str(path) + ".gz")
        assert np.array_equal(qa.csr_matrix(m).todense(), spmat.todense())

    matI = qa.Pauli("I").to_matrix()
    write_roundtrip(tmp_path / "I.mtx", matI)
    spmatI = qa.csr_matrix(matI)
    write_roundtrip2(tmp_path / "I2-gz.mtx", spmatI, symmetry="symmetric")
    write_roundtrip2(tmp_path / "I3-gz.mtx", spmatI, symmetry="hermitian")
    matY = qa.Pauli("Y").to_matrix()
    write_roundtrip(tmp_path / "Y.mtx", matY)
    write_roundtrip_gzip(tmp_path / "Y-gz.mtx", matY)
    spmatY = qa.csr_matrix(matY)
    write_roundtrip2(tmp_path / "Y2.mtx", spmatY, symmetry="skew-symmetric")
    write_roundtrip2_gzip(tmp_path / "Y2-gz.mtx", spmatY, symmetry="skew-symmetric")
    write_roundtrip2(tmp_path / "Y3.mtx", spmatY, symmetry="hermitian")
    write_roundtrip2_gzip(tmp_path / "Y3-gz.mtx", spmatY, symmetry="hermitian")


@pytest.mark.parametrize("name", ["py_H2", "py_H4", "rs_H6"])
def test_mm_text_equals_oracle(tmp_path, hamiltonians, name):
    """Written text byte-identical to the oracle's restatement, general and
    Hermitian (a Pauli sum with real coefficients is Hermitian), plain and
    gzip; reading it back gives bit-identical arrays (the floats are written
    with round-trip digits)."""
    import qrusty_amd as qa
    labels, coeffs = hamiltonians[name]
    m = _op(labels, coeffs).to_matrix_mode("Hip")
    ip, ix, dv = _host(m)
    print(f"[gpu] mm text {name}", flush=True)
    for sym in ["general", "hermitian", "symmetric", "skew-symmetric"]:
        ref = orc.mm_text(m.shape(), ip, ix, dv, sym)
        for ext in [".mtx", ".mtx.gz"]:
            path = str(tmp_path / f"{name}-{sym}{ext}")
            if sym == "general":
                m.matrixmarket_write(path)
            else:
                m.matrixmarket_write_symmetric(path, sym)
            raw = gzip.open(path).read() if ext.endswith(".gz") else open(path, "rb").read()
            assert raw.decode() == ref, (sym, ext)
    back = qa.SpMat.matrixmarket_read(str(tmp_path / f"{name}-general.mtx.gz"))
    _same(back, m)
    # the mirrored half is conj(v): equal values (an imaginary part that
    # cancelled to +0 comes back as -0)
    herm = _host(qa.SpMat.matrixmarket_read(str(tmp_path / f"{name}-hermitian.mtx")))
    assert np.array_equal(herm[0], ip) and np.array_equal(herm[1], ix)
    assert np.array_equal(herm[2], dv)


def test_mm_read_scipy_files(tmp_path):
    """Files scipy writes (every symmetry, complex and real fields, gzip) read
    into scipy's own matrix bit for bit."""
    import qrusty_amd as qa
    import scipy.io
    import scipy.sparse as sps
    rng = np.random.default_rng(12)
    n = 300
    d = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    d[rng.random((n, n)) > 0.05] = 0
    cases = {"general": d, "symmetric": d + d.T, "hermitian": d + d.conj().T,
             "skew-symmetric": d - d.T}
    for sym, mat in cases.items():
        for field_real in [False, True]:
            mm = mat.real.copy() if field_real else mat
            path = str(tmp_path / f"s-{sym}-{int(field_real)}.mtx")
            scipy.io.mmwrite(path, sps.csr_matrix(mm), symmetry=sym)
            ref = sps.csr_matrix(scipy.io.mmread(path)).astype(complex)
            ref.sort_indices()
            print(f"[gpu] mm read scipy {sym} real={field_real}", flush=True)
            got = _host(qa.SpMat.matrixmarket_read(path))
            assert np.array_equal(got[0], ref.indptr) and np.array_equal(got[1], ref.indices)
            assert np.array_equal(got[2].view(np.uint64), ref.data.view(np.uint64))
            _gzip(path)
            got2 = _host(qa.SpMat.matrixmarket_read(path + ".gz"))
            assert np.array_equal(got2[2].view(np.uint64), ref.data.view(np.uint64))


def test_mm_read_hand_written(tmp_path):
    """Comments, blank lines, CRLF, '+' signs, an integer field, duplicates
    (summed in file order), an empty matrix; and the error cases."""
    import qrusty_amd as qa
    p = tmp_path / "h.mtx"
    p.write_bytes(b"%%MatrixMarket matrix coordinate complex general\r\n% a comment\r\n\r\n"
                  b"3 4 5\r\n1 2 +0.5 -1e-3\r\n3 4 1 0\r\n1 2 0.25 0.125\r\n\r\n2 1 -0 -0\r\n"
                  b"1 2 1e-20 0\r\n")
    ip, ix, dv = _host(qa.SpMat.matrixmarket_read(str(p)))
    assert list(ip) == [0, 1, 2, 3] and list(ix) == [1, 0, 3]
    assert dv[0] == complex(((0.5 + 0.25) + 1e-20), ((-1e-3 + 0.125) + 0.0))
    assert np.signbit(dv[1].real) and np.signbit(dv[1].imag)
    p.write_text("%%MatrixMarket matrix coordinate integer symmetric\n2 2 2\n1 1 3\n2 1 -4\n")
    m = qa.SpMat.matrixmarket_read(str(p))
    assert np.array_equal(qa.csr_matrix(m).todense(), np.array([[3, -4], [-4, 0]], dtype=complex))
    p.write_text("%%MatrixMarket matrix coordinate complex general\n5 7 0\n")
    m = qa.SpMat.matrixmarket_read(str(p))
    assert m.shape() == (5, 7) and m.nnz() == 0
    bad = ["%%MatrixMarket matrix array complex general\n2 2\n1 0 0 0\n",
           "%%MatrixMarket matrix coordinate complex general\n2 2 2\n1 1 1 0\n",
           "%%MatrixMarket matrix coordinate complex general\n2 2 1\n3 1 1 0\n",
           "%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 1 x 0\n",
           "%%MatrixMarket matrix coordinate pattern general\n2 2 1\n1 1\n",
           "not a header\n"]
    for text in bad:
        p.write_text(text)
        with pytest.raises(qa.QrError):
            qa.SpMat.matrixmarket_read(str(p))
    with pytest.raises(OSError):
        qa.SpMat.matrixmarket_read(str(tmp_path / "missing.mtx"))
    with pytest.raises(OSError):
        qa.Pauli("X").to_matrix().matrixmarket_write(str(tmp_path / "no" / "dir.mtx"))
    with pytest.raises(qa.QrError, match="unrecognized mode"):
        qa.Pauli("X").to_matrix().matrixmarket_write_symmetric(str(p), "upper")
    gone = qa.Pauli("X").to_matrix()
    gone.export()
    with pytest.raises(qa.QrError, match="already-destroyed"):
        gone.matrixmarket_write(str(p))


@pytest.mark.parametrize("name", ["py_H2", "rs_H6", "py_H8"])
def test_raw_bytes_and_round_trip(tmp_path, hamiltonians, name):
    """raw_write bytes equal the oracle's rawio restatement; raw_read of
    them, of the other byte order and of gzip gives bit-identical arrays."""
    import qrusty_amd as qa
    labels, coeffs = hamiltonians[name]
    m = _op(labels, coeffs).to_matrix_mode("Hip")
    ip, ix, dv = _host(m)
    print(f"[gpu] raw {name}", flush=True)
    path = str(tmp_path / f"{name}.bin")
    m.raw_write(path)
    assert open(path, "rb").read() == orc.raw_bytes(m.shape(), ip, ix, dv, "<")
    _same(qa.SpMat.raw_read(path), m)
    swapped = str(tmp_path / f"{name}-be.bin")
    with open(swapped, "wb") as f:
        f.write(orc.raw_bytes(m.shape(), ip, ix, dv, ">"))
    _same(qa.SpMat.raw_read(swapped), m)
    m.raw_write(path + ".gz")
    assert gzip.open(path + ".gz").read() == orc.raw_bytes(m.shape(), ip, ix, dv, "<")
    _same(qa.SpMat.raw_read(path + ".gz"), m)


def test_raw_reference_and_errors(tmp_path):
    """qrusty/src/rawio.rs:216-231 (2x2 identity round trip), and malformed
    files."""
    import qrusty_amd as qa
    eye = qa.SpMat.new_unchecked((2, 2), [1 + 0j, 1 + 0j], [0, 1], [0, 1, 2])
    path = str(tmp_path / "eye.bin")
    eye.raw_write(path)
    assert np.array_equal(qa.csr_matrix(qa.SpMat.raw_read(path)).todense(), np.eye(2))
    good = open(path, "rb").read()
    for bad in [good[:1], b"XY" + good[2:], good[:2] + b"\x01" + good[3:], good[:-8]]:
        with open(path, "wb") as f:
            f.write(bad)
        with pytest.raises(qa.QrError):
            qa.SpMat.raw_read(path)
    with pytest.raises(OSError):
        qa.SpMat.raw_read(str(tmp_path / "missing.bin"))


def test_files_large(tmp_path):
    """A 19-qubit CSR (2^19 rows, about 20 M entries) through the raw
    format, and a 17-qubit one (about 5 M entries) through MatrixMarket, plain
    and gzip: all back bit-identical, compared on the device."""
    import qrusty_amd as qa
    from qrusty_amd import workloads
    for nq, nt, kinds in [(19, 40, ["raw", "raw.gz"]), (17, 40, ["mtx", "mtx.gz"])]:
        labels, coeffs = workloads.random_op(nq, nt, seed=0x10F11E + nq)
        m = qa.SparsePauliOp([qa.Pauli(l) for l in labels], coeffs).to_matrix_mode("Hip")
        a = m.device_arrays()
        for kind in kinds:
            path = str(tmp_path / f"big{nq}.{kind}")
            print(f"[gpu] files large {nq}q {kind} nnz={m.nnz()}", flush=True)
            if kind.startswith("raw"):
                m.raw_write(path)
                b = qa.SpMat.raw_read(path).device_arrays()
            else:
                m.matrixmarket_write(path)
                b = qa.SpMat.matrixmarket_read(path).device_arrays()
            for x, y in zip(a, b):
                assert torch.equal(x.view(torch.int64), y.view(torch.int64))
            os.remove(path)
