"""GPU parity of the eigensolver-step operations: SpMat.diagonal, axpby / axpy
/ ax and precond / precond2 (pyqrusty/src/lib.rs:118-125, 436-468, 519-591;
qrusty/src/accel.rs:374-393).

Bar: bit-exact against the oracle (the same num-complex formulas, no fused
multiply-add on either side).  precond compares |diag - e| with tol through
hypot, whose last bit may differ between the device and glibc, so the random
cases keep |diag - e| away from tol; the threshold itself is tested with
values whose norm is exact.  Also the reference's own tests
(pyqrusty/tests/test_it.py:46-51, 232-312), rewritten against this package.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


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


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.complex128).view(np.uint64)


@pytest.mark.parametrize("name", ["py_H2", "py_H4", "rs_H6", "py_H8"])
def test_diagonal_fixtures(hamiltonians, name):
    labels, coeffs = hamiltonians[name]
    print(f"[gpu] diagonal {name}", flush=True)
    m = _op(labels, coeffs).to_matrix_mode("Hip")
    got = m.diagonal()
    p = orc.Params(labels, coeffs)
    ip, ix, dv = orc.build_csr(p)
    assert np.array_equal(_bits(got), _bits(orc.csr_diagonal((p.N, p.N), ip, ix, dv)))
    dev = m.diagonal(device=True)
    assert dev.is_cuda and np.array_equal(_bits(dev.cpu().numpy()), _bits(got))


def test_diagonal_general_csr():
    """Rectangular shapes, missing diagonal entries, empty rows; and the
    reference's test_spmat (test_it.py:46-51)."""
    import qrusty_amd as qa
    import scipy.sparse as sps
    dmat = np.array([[1.0, 2.0], [3.0, 4.0]], dtype=complex)
    s = sps.csr_matrix(dmat)
    m = qa.SpMat.new_unchecked(dmat.shape, s.data, s.indices, s.indptr)
    assert np.array_equal(m.diagonal(), np.diagonal(dmat))
    rng = np.random.default_rng(7)
    for shape in [(3, 2), (2, 5), (64, 64), (100, 37), (1, 1)]:
        d = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        d[rng.random(shape) < 0.5] = 0.0
        d[::3, :] = 0.0  # empty rows
        s = sps.csr_matrix(d)
        s.sort_indices()
        m = qa.SpMat.new_unchecked(shape, s.data, s.indices, s.indptr)
        assert np.array_equal(m.diagonal(), np.diagonal(d)), shape
    m.export()
    with pytest.raises(qa.QrError):
        m.diagonal()


@pytest.mark.parametrize("n", [0, 1, 1000, (1 << 20) + 3])
def test_axpby_family_bit_exact(n):
    import qrusty_amd as qa
    rng = np.random.default_rng(n + 1)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    a, b = complex(0.3, -1.7), complex(-2.5, 0.25)
    print(f"[gpu] axpby n={n}", flush=True)
    assert np.array_equal(_bits(qa.axpby(a, x, b, y)), _bits(orc.axpby(a, x, b, y, 0)))
    assert np.array_equal(_bits(qa.axpy(a, x, y)), _bits(orc.axpby(a, x, 0, y, 1)))
    assert np.array_equal(_bits(qa.ax(a, x)), _bits(orc.axpby(a, x, 0, None, 2)))
    # device tensors stay on the device
    xt, yt = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    zt = qa.axpby(a, xt, b, yt)
    assert zt.is_cuda and np.array_equal(_bits(zt.cpu().numpy()), _bits(orc.axpby(a, x, b, y, 0)))


def test_reference_ax_axpy_axpby():
    """pyqrusty/tests/test_it.py:232-263."""
    import qrusty_amd as qa
    x = np.array([1.0 + 2j, 2.0 + 3j, 3.0 + 4j, 4.0 + 5j])
    y = np.array([5.0 + 6j, 6.0 + 7j, 7.0 + 8j, 8.0 + 9j])
    a, b = 0.0 + 1.0j, 2.0 - 1.0j
    assert np.allclose(qa.ax(a, x), a * x)
    assert np.allclose(qa.axpy(a, x, y), a * x + y)
    assert np.allclose(qa.axpby(a, x, b, y), a * x + b * y)
    rows = 1 << 20
    rng = np.random.default_rng(0)
    x = rng.random(rows) + rng.random(rows) * 1j
    y = rng.random(rows) + rng.random(rows) * 1j
    assert np.allclose(qa.ax(a, x), a * x)
    assert np.allclose(qa.axpby(a, x, b, y), a * x + b * y)
    with pytest.raises(qa.QrError):
        qa.axpy(a, x, y[:10])


def test_precond2_bit_exact_and_threshold():
    import qrusty_amd as qa
    rng = np.random.default_rng(11)
    n = 100000
    e, tol = complex(0.25, -0.5), 1e-3
    d = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    far = np.abs(d - e) > 2 * tol  # keep the random cases away from the threshold
    d = d[far]
    dx = rng.standard_normal(len(d)) + 1j * rng.standard_normal(len(d))
    assert np.array_equal(_bits(qa.precond2(d, dx, e, tol)), _bits(orc.precond2(d, dx, e, tol)))
    # threshold, with exact norms (e = 0, tol = 5 * 2^-12): |v| = 0 and tol/2 are
    # replaced by tol; |v| = tol (3-4-5 triangle, and -i tol) and 2 tol are not
    tol = 5.0 / 4096
    v = np.array([0.0, tol / 2, complex(3.0 / 4096, 4.0 / 4096), 2 * tol, -1j * tol])
    assert np.array_equal(np.abs(v) < tol, [True, True, False, False, False])
    dxt = rng.standard_normal(len(v)) + 1j * rng.standard_normal(len(v))
    got = qa.precond2(v, dxt, 0.0, tol)
    ref = orc.precond2(v, dxt, 0.0, tol)
    assert np.array_equal(_bits(got), _bits(ref))
    assert abs(got[0] - dxt[0] / tol) <= 2e-15 * abs(dxt[0] / tol)
    assert abs(got[4] - dxt[4] / (-1j * tol)) <= 2e-15 * abs(dxt[4] / tol)


def test_precond_from_csr(hamiltonians):
    """precond reads the diagonal from the CSR: equals precond2 on
    SpMat.diagonal(), and the reference's test_precond / test_precond2 and
    printed vector (test_it.py:276-312)."""
    import qrusty_amd as qa
    with open(os.path.join(os.path.dirname(__file__), "golden", "precond_h2.json")) as f:
        pc = json.load(f)
    dx = np.array([complex(a, b) for a, b in pc["dx"]])
    rv = np.array([complex(a, b) for a, b in pc["rv_printed"]])
    labels, coeffs = hamiltonians["py_H2"]
    m = _op(labels, coeffs).to_matrix()
    r1 = qa.precond(m, dx, pc["e"], pc["tol"])
    r2 = qa.precond2(m.diagonal(), dx, pc["e"], pc["tol"])
    assert np.array_equal(_bits(r1), _bits(r2))
    diag = m.diagonal()

    def reg(v, tol):
        return np.where(np.abs(v) < tol, tol, v)
    assert np.allclose(r1, dx / reg(diag - pc["e"], pc["tol"]))
    assert np.allclose(r1, rv, rtol=1e-8, atol=1e-20)
    labels, coeffs = hamiltonians["py_H8"]
    m = _op(labels, coeffs).to_matrix_mode("Hip")
    dx = np.random.default_rng(3).standard_normal(1 << 16) + 0j
    assert np.array_equal(_bits(qa.precond(m, dx, -3.0, 1e-14)),
                          _bits(qa.precond2(m.diagonal(), dx, -3.0, 1e-14)))
    m.export()
    with pytest.raises(qa.QrError):
        qa.precond(m, dx, -3.0, 1e-14)


def _sp(qa, d):
    import scipy.sparse as sps
    s = sps.csr_matrix(d)
    s.sort_indices()
    return s, qa.SpMat.new_unchecked(d.shape, s.data, s.indices, s.indptr)


def _host(m):
    _, data, indices, indptr = m.__copy__().export()
    return indptr, indices, data


def test_spadd_bit_exact_random():
    """a*A + b*B against the oracle's merge: indptr and indices equal, values
    to the last bit; exact cancellations vanish; rectangular shapes, empty
    rows, one side empty."""
    import qrusty_amd as qa
    rng = np.random.default_rng(21)
    # the last two: rows longer than a staged tile (the global-memory merge),
    # and mostly short rows with a few long ones (both routes in one launch)
    for shape, dens in [((200, 200), 0.05), ((37, 300), 0.2), ((1, 1), 1.0), ((64, 5), 0.0),
                        ((5000, 5000), 0.002), ((6, 3000), 0.3), ((3000, 2000), "uneven")]:
        if dens == "uneven":
            dens = np.where(np.arange(shape[0])[:, None] % 97 == 5, 0.5, 0.002)
        d1 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        d2 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        d1[rng.random(shape) >= dens] = 0.0
        d2[rng.random(shape) >= np.minimum(2 * np.asarray(dens), 1.0)] = 0.0
        d1[::4, :] = 0.0
        # -0.0 components: one-sided entries get + 0, which makes them +0.0
        d1[1::4, :] = np.where(d1[1::4, :] != 0, np.abs(d1[1::4, :].real) * -1 - 0.0j, 0)
        d1[1::4, :].imag[d1[1::4, :] != 0] = -0.0
        k = min(shape) // 2
        d2[:k, :k] = -d1[:k, :k]  # cancel exactly for a = b = 1
        s1, m1 = _sp(qa, d1)
        s2, m2 = _sp(qa, d2)
        for a, b in [(1.0, 1.0), (0.3 - 1.1j, -2.0 + 0.5j), (0.0, 1j)]:
            print(f"[gpu] spadd shape={shape} a={a} b={b}", flush=True)
            got = _host(qa.a_spmat_p_b_spmat(a, m1, b, m2))
            ref = orc.csr_add(shape[0], a, (s1.indptr, s1.indices, s1.data), b,
                              (s2.indptr, s2.indices, s2.data))
            assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
            assert np.array_equal(_bits(got[2]), _bits(ref[2]))


def test_spadd_malformed_stays_in_bounds():
    """A CSR handed in through new_unchecked with an indptr that descends in
    places and runs past nnz: the count and fill passes tile the rows
    differently and may clamp such rows differently (the entries of such rows
    are then unspecified), but every read and write stays in bounds: the call
    returns, with an indptr that ascends from 0 to the output's entry count."""
    import qrusty_amd as qa
    rng = np.random.default_rng(77)
    rows, cols = 4000, 3000
    lens = rng.integers(0, 60, size=rows)
    lens[rng.choice(rows, size=40, replace=False)] = rng.integers(300, 1200, size=40)
    indptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    nnz = int(indptr[-1])
    indices = np.concatenate([np.sort(rng.choice(cols, size=int(n), replace=False)) for n in lens]
                             + [np.zeros(0, dtype=np.int64)]).astype(np.uint64)
    data = rng.standard_normal(nnz) + 1j * rng.standard_normal(nnz)
    bad = indptr.copy()
    swap = rng.choice(np.arange(1, rows), size=200, replace=False)
    bad[swap] = bad[swap] + np.uint64(5000)  # descends at the next row
    bad[rows - 30:] += np.uint64(10**7)      # the last rows start beyond nnz
    m_bad = qa.SpMat.new_unchecked((rows, cols), data, indices, bad)
    m_ok = qa.SpMat.new_unchecked((rows, cols), data, indices, indptr)
    print("[gpu] spadd malformed indptr", flush=True)
    for a, b in [(1.0, 1.0), (0.5, -2.0)]:
        ip, ix, dv = _host(qa.a_spmat_p_b_spmat(a, m_bad, b, m_ok))
        assert ip[0] == 0 and np.all(np.diff(ip.astype(np.int64)) >= 0)
        assert int(ip[-1]) == ix.size == dv.size


def test_spadd_pauli_shift(hamiltonians):
    """H - e*I on a Pauli CSR (the Davidson shift): bit-exact vs the oracle and
    equal to the dense a*H + b*I; SpMat + and - and scale."""
    import qrusty_amd as qa
    labels, coeffs = hamiltonians["py_H8"]
    h = _op(labels, coeffs).to_matrix_mode("Hip")
    n = h.shape()[0]
    eye = qa.SparsePauliOp([qa.Pauli("I" * len(labels[0]))], [1.0]).to_matrix()
    e = -3.25 + 0.125j
    print("[gpu] spadd H8 shift", flush=True)
    got = _host(qa.a_spmat_p_b_spmat(1.0, h, -e, eye))
    p = orc.Params(labels, coeffs)
    ref = orc.csr_add(n, 1.0, orc.build_csr(p), -e,
                      (np.arange(n + 1, dtype=np.uint64), np.arange(n, dtype=np.uint64),
                       np.ones(n, complex)))
    for g_, r_ in zip(got, ref):
        assert np.array_equal(np.asarray(g_).view(np.uint64), np.asarray(r_).view(np.uint64))
    hs = _host(h)
    s = _host(h + eye)
    d = _host(h - eye)
    r_add = orc.csr_add(n, 1.0, hs, 1.0, _host(eye))
    r_sub = orc.csr_add(n, 1.0, hs, -1.0, _host(eye))
    for g_, r_ in list(zip(s, r_add)) + list(zip(d, r_sub)):
        assert np.array_equal(np.asarray(g_).view(np.uint64), np.asarray(r_).view(np.uint64))
    z = _host(h - h)
    assert z[1].size == 0 and np.array_equal(z[0], np.zeros(n + 1, dtype=np.uint64))
    hc = h.__copy__()
    hc.scale(0.5 - 2j)
    assert np.array_equal(_bits(_host(hc)[2]), _bits(orc.axpby(0.5 - 2j, hs[2], 0, None, 2)))


def test_reference_add_interpolate():
    """pyqrusty/tests/test_it.py:103-139."""
    import qrusty_amd as qa
    import scipy.sparse as sps  # noqa: F401
    I = np.array([[1, 0], [0, 1]], dtype=complex)
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    matI = qa.Pauli("I").to_matrix()
    matX = qa.Pauli("X").to_matrix()
    assert np.array_equal(qa.csr_matrix(matI + matX).todense(), I + X)
    a, b = complex(1.0), complex(1.0j)
    assert np.array_equal(qa.csr_matrix(qa.a_spmat_p_b_spmat(a, matI, b, matX)).todense(),
                          a * I + b * X)
    with pytest.raises(Exception):
        qa.a_spmat_p_b_spmat(a, matI, b, qa.Pauli("XI").to_matrix())
    assert np.array_equal(qa.csr_matrix(matI - matI).todense(), I - I)
    gone = qa.Pauli("X").to_matrix()
    gone.export()
    with pytest.raises(qa.QrError, match="first argument"):
        qa.a_spmat_p_b_spmat(a, gone, b, matX)
    with pytest.raises(qa.QrError, match="second argument"):
        qa.a_spmat_p_b_spmat(a, matX, b, gone)
    with pytest.raises(qa.QrError):
        gone + matX


def test_reference_nz():
    """pyqrusty/tests/test_it.py:140-147."""
    import qrusty_amd as qa
    matI = qa.Pauli("I").to_matrix()
    matI.scale(1e-8 + 0j)
    assert matI.count_zeros() == 2
    mat2 = matI.eliminate_zeros()
    assert mat2.count_zeros() == 0
    assert mat2.nnz() == 0
    gone = qa.Pauli("I").to_matrix()
    gone.export()
    with pytest.raises(qa.QrError, match="count zeroes"):
        gone.count_zeros()
    with pytest.raises(qa.QrError, match="eliminate zeroes"):
        gone.eliminate_zeros()


@pytest.mark.parametrize("tol", [0.0, 0.05, 0.5, 5.0])
def test_zero_filters_vs_oracle(tol):
    """count_zeros / eliminate_zeros bit-exact vs the oracle on a random CSR
    (magnitudes kept away from tol, where device and host hypot could round
    apart) and on |3 + 4i| = 5 exactly at the threshold."""
    import qrusty_amd as qa
    import scipy.sparse as sps
    rng = np.random.default_rng(int(tol * 100) + 3)
    m = sps.random(3000, 700, density=0.05, random_state=9, format="csr", dtype=np.float64)
    mag = m.data * 2
    mag = np.where(np.abs(mag - tol) < 1e-6, mag + 1e-3, mag)
    m = sps.csr_matrix((mag * np.exp(2j * np.pi * rng.random(m.nnz)), m.indices, m.indptr),
                       shape=m.shape)
    m.sort_indices()
    m.data[::11] = 3 + 4j
    sm = qa.SpMat.new_unchecked(m.shape, m.data, m.indices, m.indptr)
    print(f"[gpu] zero filters tol={tol}", flush=True)
    assert sm.count_zeros(tol) == orc.count_zeros(m.data, tol)
    got = _host(sm.eliminate_zeros(tol))
    ref = orc.eliminate_zeros(m.shape[0], (m.indptr, m.indices, m.data), tol)
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    assert np.array_equal(_bits(got[2]), _bits(ref[2]))
    five = qa.SpMat.new_unchecked((1, 1), [3 + 4j], [0], [0, 1])
    assert five.count_zeros(5.0) == 1 and five.count_zeros(np.nextafter(5.0, 0)) == 0
    assert five.eliminate_zeros(5.0).nnz() == 0 and five.eliminate_zeros(4.99).nnz() == 1


def test_device_handoff_zero_copy():
    """Device tensors in and out without a copy: new_unchecked adopts torch
    tensors, device_arrays returns the same storage, DLPack round trip."""
    import qrusty_amd as qa
    from torch.utils import dlpack
    m = qa.SparsePauliOp([qa.Pauli("XZ"), qa.Pauli("YY")], [1.0, 0.5j]).to_matrix()
    ip, ix, dv = m.device_arrays()
    m2 = qa.SpMat.new_unchecked(m.shape(), dlpack.from_dlpack(dlpack.to_dlpack(dv)), ix, ip)
    assert m2.device_arrays()[2].data_ptr() == dv.data_ptr()
    x = np.arange(4) + 1j
    assert np.array_equal(qa.spmat_dot_densevec(m2, x), qa.spmat_dot_densevec(m, x))
    with pytest.raises(qa.QrError):
        qa.SpMat.new_unchecked(m.shape(), dv.to(torch.complex64), ix, ip)
