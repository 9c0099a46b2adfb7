"""GPU parity of the HIP CSR build and CSR matvec against the oracle.

Bar: indptr/indices/data bit-exact (SHA-256 of the exported bytes, or byte
equality with the oracle's rowwise build); matvec within 1e-12 relative
(normwise, and componentwise against sum_k |a_k||x_k|).  Also the reference's
own Python tests for this path (pyqrusty/tests/test_H.py, test_it.py:28-230),
rewritten against this package.
"""
import hashlib

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _op(labels, coeffs):
    import qrusty_amd as qa
    return qa.SparsePauliOp([qa.Pauli(l) for l in labels], coeffs)


def _export(spmat):
    shape, data, indices, indptr = spmat.export()
    return shape, data, indices, indptr


def _check_matvec(y, ref, A_abs_x):
    err = np.abs(y - ref)
    assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)
    assert np.all(err <= 1e-12 * A_abs_x + 1e-300)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import qrusty_amd  # noqa: F401  (fails loudly if libqrhip.so is missing)
    yield
    torch.cuda.synchronize()


@pytest.mark.parametrize("name", ["py_H2", "rs_H2", "py_H4", "rs_H4", "py_H6", "rs_H6", "py_H8"])
def test_fixture_csr_bit_exact(hamiltonians, csr_digests, name):
    print(f"[gpu] csr fixture {name}", flush=True)
    labels, coeffs = hamiltonians[name]
    m = _op(labels, coeffs).to_matrix_mode("RowwiseUnsafeChunked/100")
    shape, data, indices, indptr = _export(m)
    dg = csr_digests[name]
    assert shape == (1 << dg["num_qubits"],) * 2
    assert len(data) == dg["nnz"]
    assert _sha(indptr) == dg["indptr_sha256"]
    assert _sha(indices) == dg["indices_sha256"]
    assert _sha(data) == dg["data_sha256"]


def test_config1_build_and_matvec(config1):
    print("[gpu] config1 12q/200", flush=True)
    import qrusty_amd as qa
    op = _op(config1["labels"], config1["coeffs"])
    m = op.to_matrix_mode("RowwiseUnsafeChunked/100")
    from qrusty_amd import workloads
    x = workloads.random_states(1 << 12, 1, seed=workloads.SEED_CFG1 + 1)[0]
    y = qa.spmat_dot_densevec(m, x)
    shape, data, indices, indptr = _export(m)
    assert _sha(indices) == config1["csr"]["indices_sha256"]
    assert _sha(data) == config1["csr"]["data_sha256"]
    assert _sha(indptr) == config1["csr"]["indptr_sha256"]
    import scipy.sparse as sps
    A = sps.csr_matrix((np.abs(data), indices.astype(np.int64), indptr.astype(np.int64)), shape=shape)
    _check_matvec(y, config1["y"], A.dot(np.abs(x)))


def _random_labels(rng, n, T, prefixes=True):
    pre = ["", "+", "-", "i", "j", "-i", "+j", "1", "-1", "-1j"] if prefixes else [""]
    labels = [str(rng.choice(pre)) + "".join(rng.choice(list("IXYZ"), size=n)) for _ in range(T)]
    for _ in range(T // 4):
        labels[rng.integers(0, T)] = labels[rng.integers(0, T)]
    return labels


@pytest.mark.parametrize("seed", range(8))
def test_random_ops_bit_exact(seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.integers(1, 13))
    T = int(rng.integers(1, 300))
    labels = _random_labels(rng, n, T)
    coeffs = rng.standard_normal(T) + 1j * rng.standard_normal(T)
    coeffs[rng.random(T) < 0.1] = 0.0
    print(f"[gpu] random op seed={seed} n={n} T={T}", flush=True)
    p = orc.Params(labels, coeffs)
    ip, ind, dat = orc.build_csr(p)
    shape, data, indices, indptr = _export(_op(labels, coeffs).to_matrix_mode("Hip"))
    assert indptr.tobytes() == ip.tobytes()
    assert indices.tobytes() == ind.tobytes()
    assert data.tobytes() == dat.tobytes()


def test_kron_modes_match_dense():
    """Modes "", Binary, Reduce, Rayon follow Pauli::to_matrix's i^base_phase
    (lib.rs:182-213, 401-412)."""
    rng = np.random.default_rng(9)
    labels = _random_labels(rng, 5, 40)
    coeffs = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    dense_k = orc.dense_sum(labels, coeffs, "kron")
    dense_r = orc.dense_sum(labels, coeffs, "rowwise")
    import qrusty_amd as qa
    op = _op(labels, coeffs)
    for mode, ref in [("", dense_k), ("Binary", dense_k), ("Reduce", dense_k), ("Rayon", dense_k),
                      ("Accel", dense_r), ("Rowwise", dense_r), ("RowwiseUnsafe", dense_r),
                      ("RayonChunked/7", dense_r), ("RowwiseUnsafeChunked/3", dense_r)]:
        got = qa.csr_matrix(op.to_matrix_mode(mode)).toarray()
        assert np.array_equal(got, ref), mode


def test_build_rows_slices_and_wide_masks():
    """qr_csr_build_rows on row slices, including 33- to 40-qubit operators (64-bit
    masks), for every build route: 1 and 2 group chunks per thread (G <= 1024,
    2048) and sub-operators (G > 2048: two to five subtrees of the mask trie,
    with 32- and 64-bit masks)."""
    import ctypes
    import qrusty_amd as qa
    from qrusty_amd import _lib as L
    rng = np.random.default_rng(21)
    for n, T in [(10, 50), (33, 20), (40, 64), (12, 1500), (13, 2600), (13, 7000), (14, 9000),
                 (36, 5000), (36, 7000)]:
        labels = _random_labels(rng, n, T, prefixes=False)
        coeffs = rng.standard_normal(T) + 1j * rng.standard_normal(T)
        print(f"[gpu] build_rows n={n} T={T}", flush=True)
        op = _op(labels, coeffs)
        h = op._handle(L.QR_PHASE_ROWWISE)
        p = orc.Params(labels, coeffs)
        N = 1 << n
        for r0 in [0, N // 3, N - 300]:
            r1 = r0 + 300
            ind = torch.empty(300 * h.G, dtype=torch.int64, device="cuda")
            dat = torch.empty(300 * h.G, dtype=torch.complex128, device="cuda")
            L.check(L.lib.qr_csr_build_rows(h.h, r0, r1, ctypes.c_void_p(ind.data_ptr()),
                                            ctypes.c_void_p(dat.data_ptr()), None))
            torch.cuda.synchronize()
            _, ri, rd = orc.build_csr(p, row_begin=r0, row_end=r1, step=64)
            assert ind.cpu().numpy().view(np.uint64).tobytes() == ri.tobytes()
            assert dat.cpu().numpy().tobytes() == rd.tobytes()


def test_single_term_and_identity_ops():
    for labels in [["I"], ["Y"], ["Z" * 9], ["X" * 11], ["I" * 12, "I" * 12]]:
        coeffs = [1.5 - 0.25j] * len(labels)
        p = orc.Params(labels, coeffs)
        ip, ind, dat = orc.build_csr(p)
        _, data, indices, indptr = _export(_op(labels, coeffs).to_matrix_mode("Hip"))
        assert indices.tobytes() == ind.tobytes() and data.tobytes() == dat.tobytes()


# ---------------------------------------------------------------------------
# the reference's Python tests for this path
# ---------------------------------------------------------------------------
I2 = np.array([[1.0, 0.0], [0.0, 1.0]], dtype=complex)
Z2 = np.array([[1.0, 0.0], [0.0, -1.0]], dtype=complex)
X2 = np.array([[0.0, 1.0], [1.0, 0.0]], dtype=complex)
Y2 = np.array([[0.0, -1.0j], [1.0j, 0.0]], dtype=complex)


def test_reference_test_H(hamiltonians):
    """pyqrusty/tests/test_H.py:28-52."""
    import qrusty_amd as qa
    for name in ["py_H2", "py_H4", "py_H6"]:
        H = _op(*hamiltonians[name])
        m1 = qa.csr_matrix(H.to_matrix())
        assert np.array_equal(m1.todense(), qa.csr_matrix(H.to_matrix_mode(mode="")).todense())
        m2 = H.to_matrix_mode(mode="RowwiseUnsafeChunked/100")
        assert np.array_equal(m1.todense(), qa.csr_matrix(m2).todense())
    with pytest.raises(Exception):
        _op(*hamiltonians["py_H2"]).to_matrix_mode("foo")
    assert _op(*hamiltonians["py_H6"]).to_matrix().shape() == (1 << 12, 1 << 12)


def test_reference_test_it_paulis_and_spmat():
    """pyqrusty/tests/test_it.py:46-101, 218-230."""
    import scipy.sparse as sps
    import qrusty_amd as qa
    for lab, M in [("I", I2), ("X", X2), ("Z", Z2), ("Y", Y2)]:
        assert np.array_equal(M, qa.csr_matrix(qa.Pauli(lab).to_matrix()).todense())
    spop = qa.SparsePauliOp([qa.Pauli("I"), qa.Pauli("Y")], [1.0 + 0.0j, 2.0 + 0.0j])
    assert np.array_equal(I2 + 2.0 * Y2, qa.csr_matrix(spop.to_matrix()).todense())

    dmat = np.array([[1.0, 2.0], [3.0, 4.0]], dtype=complex)
    spmat = sps.csr_matrix(dmat)
    spmat2 = qa.SpMat.new_unchecked(dmat.shape, spmat.data, spmat.indices, spmat.indptr)
    assert repr(spmat2) == ("<2x2 sparse matrix of type Complex64\n\twith 4 stored elements in "
                            "Compressed Sparse Row format>")
    assert np.array_equal(spmat2.diagonal(), np.diagonal(dmat))
    dmat3 = np.array([[1.0, 2.0], [3.0, 4.0], [5.0, 6.0]], dtype=complex)
    s3 = sps.csr_matrix(dmat3)
    spmat3 = qa.SpMat.new_unchecked(s3.shape, s3.data, s3.indices, s3.indptr)
    assert repr(spmat3) == ("<3x2 sparse matrix of type Complex64\n\twith 6 stored elements in "
                            "Compressed Sparse Row format>")
    x2 = np.array([1.0 + 1j, -2.0 + 0.5j])
    assert np.allclose(qa.spmat_dot_densevec(spmat3, x2), dmat3 @ x2, rtol=1e-15, atol=0)
    qa.csr_matrix(spmat3)
    with pytest.raises(Exception):
        spmat3.export()
    with pytest.raises(Exception):
        qa.spmat_dot_densevec(spmat3, x2)
    assert repr(spmat3) == "<already-dropped sparse matrix of type Complex64>"

    pIX = qa.Pauli("IX")
    matIX = pIX.to_matrix()
    csr_matIX = qa.csr_matrix(pIX.to_matrix())
    x = np.array([1.0 + 2j, 2.0 + 3j, 3.0 + 4j, 4.0 + 5j])
    assert np.allclose(qa.spmat_dot_densevec(matIX, x), csr_matIX.dot(x))
    with pytest.raises(Exception):
        qa.spmat_dot_densevec(matIX, x[:3])


def test_matvec_fixtures_vs_oracle(hamiltonians):
    import scipy.sparse as sps
    import qrusty_amd as qa
    rng = np.random.default_rng(17)
    for name in ["py_H6", "rs_H6", "py_H8"]:
        labels, coeffs = hamiltonians[name]
        print(f"[gpu] matvec {name}", flush=True)
        m = _op(labels, coeffs).to_matrix_mode("Hip")
        n = m.shape()[0]
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        y = qa.spmat_dot_densevec(m, x)
        shape, data, indices, indptr = _export(m)
        ref = orc.spmat_dot_densevec(indptr, indices, data, x)
        A = sps.csr_matrix((np.abs(data), indices.astype(np.int64), indptr.astype(np.int64)), shape=shape)
        _check_matvec(y, ref, A.dot(np.abs(x)))


def test_matvec_general_csr_shapes():
    """Arbitrary CSR (empty rows, ragged rows, rectangular) through new_unchecked."""
    import scipy.sparse as sps
    import qrusty_amd as qa
    rng = np.random.default_rng(5)
    for rows, cols, dens in [(1, 1, 1.0), (7, 3, 0.5), (1000, 200, 0.01), (300, 5000, 0.2),
                             (64, 64, 0.0)]:
        A = sps.random(rows, cols, density=dens, format="csr", random_state=3,
                       dtype=np.complex128,
                       data_rvs=lambda k: rng.standard_normal(k) + 1j * rng.standard_normal(k))
        A.sort_indices()
        m = qa.SpMat.new_unchecked(A.shape, A.data, A.indices, A.indptr)
        x = rng.standard_normal(cols) + 1j * rng.standard_normal(cols)
        y = qa.spmat_dot_densevec(m, x)
        ref = orc.spmat_dot_densevec(A.indptr.astype(np.uint64), A.indices.astype(np.uint64), A.data, x)
        _check_matvec(y, ref, abs(A).dot(np.abs(x)))


def test_matvec_short_rows_ragged():
    """16..32 entries per row on average (the two-rows-per-slot kernel) with
    empty rows and rows of up to 300 entries, and a row count that is not a
    multiple of the 64 rows a block covers."""
    import scipy.sparse as sps
    import qrusty_amd as qa
    rng = np.random.default_rng(11)
    for rows, cols in [(3001, 500), (64, 40), (130, 4000)]:
        lens = rng.integers(0, 41, size=rows)
        lens[rng.choice(rows, size=max(1, rows // 50), replace=False)] = rng.integers(33, 301)
        lens[rng.choice(rows, size=max(1, rows // 20), replace=False)] = 0
        lens = np.minimum(lens, cols)
        indptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        indices = np.concatenate([np.sort(rng.choice(cols, size=int(n), replace=False)) for n in lens]
                                 + [np.zeros(0, dtype=np.int64)]).astype(np.int64)
        data = rng.standard_normal(indptr[-1]) + 1j * rng.standard_normal(indptr[-1])
        A = sps.csr_matrix((data, indices, indptr), shape=(rows, cols))
        avg = indptr[-1] // rows
        print(f"[gpu] short-row matvec rows={rows} cols={cols} avg={avg}", flush=True)
        m = qa.SpMat.new_unchecked(A.shape, A.data, A.indices, A.indptr)
        x = rng.standard_normal(cols) + 1j * rng.standard_normal(cols)
        y = qa.spmat_dot_densevec(m, x)
        ref = orc.spmat_dot_densevec(A.indptr.astype(np.uint64), A.indices.astype(np.uint64), A.data, x)
        _check_matvec(y, ref, abs(A).dot(np.abs(x)))


@pytest.mark.parametrize("rows,cols,unsorted", [(1, 3000, False), (7, 700, True), (9, 5000, False),
                                                (33, 2048, True), (1001, 1500, False),
                                                (4096, 1 << 14, False)])
def test_matvec_long_rows_ragged(rows, cols, unsorted):
    """256+ entries per row on average (the kernel that stages each wave's 8
    rows in LDS): rows of 0 to 3000 entries side by side in one wave, chunk
    tails, row counts that are not a multiple of 8 rows per wave or 32 per
    block, unsorted columns and repeated columns within a row."""
    import scipy.sparse as sps
    import qrusty_amd as qa
    rng = np.random.default_rng(rows * 7 + cols)
    lens = rng.integers(256, 700, size=rows)
    lens[rng.choice(rows, size=max(1, rows // 9), replace=False)] = rng.integers(0, 3001)
    lens[rng.choice(rows, size=max(1, rows // 13), replace=False)] = 0
    lens[rng.choice(rows, size=max(1, rows // 11), replace=False)] = 65
    indptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    parts = []
    for n in lens:
        c = rng.integers(0, cols, size=int(n))
        parts.append(c if unsorted else np.sort(c))
    indices = np.concatenate(parts + [np.zeros(0, dtype=np.int64)]).astype(np.int64)
    data = rng.standard_normal(indptr[-1]) + 1j * rng.standard_normal(indptr[-1])
    avg = indptr[-1] // rows
    print(f"[gpu] long-row matvec rows={rows} cols={cols} avg={avg} unsorted={unsorted}", flush=True)
    assert avg >= 256 or rows < 8
    m = qa.SpMat.new_unchecked((rows, cols), data, indices, indptr)
    x = rng.standard_normal(cols) + 1j * rng.standard_normal(cols)
    y = qa.spmat_dot_densevec(m, x)
    ref = orc.spmat_dot_densevec(indptr.astype(np.uint64), indices.astype(np.uint64), data, x)
    A = sps.csr_matrix((np.abs(data), indices, indptr), shape=(rows, cols))
    _check_matvec(y, ref, A.dot(np.abs(x)))


@pytest.mark.parametrize("avg", [20, 100, 400])
def test_matvec_malformed_csr_skips_out_of_range(avg):
    """A CSR handed in through new_unchecked with columns >= ncols (some near
    2^64) and indptr entries beyond nnz: every matvec kernel (16-32, fewer
    than 256 and 256+ entries per row on average) skips those columns and
    clamps the rows to [0, nnz), so y is the sum over the remaining entries."""
    import qrusty_amd as qa
    rng = np.random.default_rng(900 + avg)
    rows, cols = 333, 5000
    lens = rng.integers(avg // 2, avg + avg // 2 + 1, size=rows)
    indptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    nnz = int(indptr[-1])
    indices = rng.integers(0, cols, size=nnz).astype(np.uint64)
    bad = rng.random(nnz) < 0.05
    indices[bad] = cols + rng.integers(0, 1000, size=int(bad.sum())).astype(np.uint64)
    indices[rng.random(nnz) < 0.01] = np.uint64(2**64 - 7)
    indptr[rows - 20:] += np.uint64(10**6)  # the last rows end (and start) beyond nnz
    data = rng.standard_normal(nnz) + 1j * rng.standard_normal(nnz)
    x = rng.standard_normal(cols) + 1j * rng.standard_normal(cols)
    print(f"[gpu] malformed CSR avg={avg} nnz={nnz}", flush=True)
    m = qa.SpMat.new_unchecked((rows, cols), data, indices, indptr)
    y = qa.spmat_dot_densevec(m, x)
    ref = np.zeros(rows, dtype=np.complex128)
    absb = np.zeros(rows)
    for r in range(rows):
        k0, k1 = min(int(indptr[r]), nnz), min(int(indptr[r + 1]), nnz)
        sl = slice(k0, max(k0, k1))
        ok = indices[sl] < cols
        c = indices[sl][ok].astype(np.int64)
        ref[r] = np.sum(data[sl][ok] * x[c])
        absb[r] = np.sum(np.abs(data[sl][ok]) * np.abs(x[c]))
    assert np.all(np.abs(y - ref) <= 1e-12 * absb + 1e-300)


def test_device_tensor_matvec_stays_on_device(hamiltonians):
    import qrusty_amd as qa
    m = _op(*hamiltonians["py_H4"]).to_matrix_mode("Hip")
    x = torch.randn(256, dtype=torch.complex128, device="cuda")
    y = qa.spmat_dot_densevec(m, x)
    assert y.is_cuda and y.shape == (256,)
    ref = qa.spmat_dot_densevec(m, x.cpu().numpy())
    assert np.array_equal(y.cpu().numpy(), ref)



@pytest.mark.parametrize("n", [14, 34])
def test_sub_operators_with_heavy_group(n):
    """More than 2048 groups (sub-operators) plus a diagonal group of 150 Z
    strings (a heavy group, folded once per row): the CSR equals the oracle's
    rows bit for bit, with 32- and 64-bit masks."""
    import qrusty_amd as qa
    rng = np.random.default_rng(77 + n)
    labels = _random_labels(rng, n, 3200, prefixes=False)
    diag = ["".join(rng.choice(["I", "Z"], size=n)) for _ in range(150)]
    labels = labels + diag
    coeffs = rng.standard_normal(len(labels)) + 1j * rng.standard_normal(len(labels))
    p = orc.Params(labels, coeffs)
    G = len(np.unique(p.x))
    print(f"[gpu] sub-operators + heavy n={n} G={G}", flush=True)
    assert G > 2048
    op = _op(labels, coeffs)
    N = 1 << n
    rows = np.unique(np.concatenate([[0, 1, 2, 3, N - 1], rng.integers(0, N, 40)])).astype(np.uint64)
    if n <= 14:
        shape, data, indices, indptr = _export(op.to_matrix_mode("Hip"))
        ip, ind, dat = orc.build_csr(p)
        assert indptr.tobytes() == ip.tobytes() and indices.tobytes() == ind.tobytes()
        assert data.view(np.uint64).tobytes() == dat.view(np.uint64).tobytes()
    else:
        import ctypes
        from qrusty_amd import _lib as L
        h = op._handle(0)
        ix = torch.empty(G * 16, dtype=torch.int64, device="cuda")
        dv = torch.empty(G * 16, dtype=torch.complex128, device="cuda")
        st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        for r in rows:
            r0 = int(r) & ~15
            L.check(L.lib.qr_csr_build_rows(h.h, r0, r0 + 16, ctypes.c_void_p(ix.data_ptr()),
                                            ctypes.c_void_p(dv.data_ptr()), st))
            got_ix = ix.cpu().numpy().view(np.uint64).reshape(16, G)[int(r) - r0]
            got_dv = dv.cpu().numpy().reshape(16, G)[int(r) - r0]
            cols, vals = orc.make_row(p, int(r))
            assert np.array_equal(got_ix, cols), f"row {r}"
            assert np.array_equal(got_dv.view(np.uint64), vals.view(np.uint64)), f"row {r}"


def _grouped_labels(rng, n, sizes):
    """Labels in groups: one random x mask per entry of `sizes`, with that many
    terms of random z masks (Y where x and z meet)."""
    masks = rng.choice(1 << n, size=len(sizes), replace=False)
    chars = {(0, 0): "I", (0, 1): "Z", (1, 0): "X", (1, 1): "Y"}
    labels = []
    for xm, k in zip(masks, sizes):
        for zm in rng.integers(0, 1 << n, size=k):
            labels.append("".join(chars[((int(xm) >> q) & 1, (int(zm) >> q) & 1)]
                                  for q in reversed(range(n))))
    return labels


@pytest.mark.parametrize("case", ["single", "sub"])
def test_tail_table_groups_bit_exact(case):
    """Operators with heavy groups fold the terms past each group's second
    from a transposed tail table (at most 16 positions).  Groups of 6 terms
    (the molecular shape), groups of 20 (tails past the table's 16
    positions, read from the group's own terms), and more heavy groups than
    one tile's LDS holds (the ones left out run the batch loop, past the
    table) -- single operator (two-chunk heavy instance) and sub-operators
    (G > 2048, the table offset per subtree).  All rows bit-exact."""
    rng = np.random.default_rng(4711 if case == "single" else 4712)
    if case == "single":
        n, sizes = 11, [6] * 150 + [20] * 3 + [60] * 20
    else:
        n, sizes = 13, [4] * 2100 + [6] * 40 + [60] * 10
    labels = _grouped_labels(rng, n, sizes)
    order = rng.permutation(len(labels))
    labels = [labels[i] for i in order]
    coeffs = rng.standard_normal(len(labels)) + 1j * rng.standard_normal(len(labels))
    p = orc.Params(labels, coeffs)
    G = len(np.unique(p.x))
    print(f"[gpu] tail table {case} n={n} T={len(labels)} G={G}", flush=True)
    assert G == len(sizes) and (G > 2048) == (case == "sub")
    ip, ind, dat = orc.build_csr(p)
    shape, data, indices, indptr = _export(_op(labels, coeffs).to_matrix_mode("Hip"))
    assert indptr.tobytes() == ip.tobytes()
    assert indices.tobytes() == ind.tobytes()
    assert data.view(np.uint64).tobytes() == dat.view(np.uint64).tobytes()


def test_heavy_fold_row_windows_bit_exact():
    """qr_csr_build_rows over row windows that start and end inside a tile,
    for an operator whose two-chunk heavy-group instance builds several
    tiles per block with one heavy fold: each window's rows equal the
    oracle's bit for bit (first tile unaligned, last block partial, rows past
    the window inside a fold)."""
    import ctypes
    from qrusty_amd import _lib as L
    rng = np.random.default_rng(4713)
    n, sizes = 11, [6] * 150 + [20] * 3 + [60] * 20
    labels = _grouped_labels(rng, n, sizes)
    coeffs = rng.standard_normal(len(labels)) + 1j * rng.standard_normal(len(labels))
    p = orc.Params(labels, coeffs)
    G = len(sizes)
    ip, ind, dat = orc.build_csr(p)
    ind = ind.reshape(-1, G)
    dat = dat.view(np.uint64).reshape(-1, 2 * G)
    h = _op(labels, coeffs)._handle(0)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    N = 1 << n
    for r0, r1 in [(5, 300), (129, 130), (1000, N - 3), (0, N), (N - 77, N)]:
        print(f"[gpu] heavy fold rows [{r0}, {r1})", flush=True)
        W = r1 - r0
        ix = torch.empty(W * G, dtype=torch.int64, device="cuda")
        dv = torch.empty(W * G, dtype=torch.complex128, device="cuda")
        L.check(L.lib.qr_csr_build_rows(h.h, r0, r1, ctypes.c_void_p(ix.data_ptr()),
                                        ctypes.c_void_p(dv.data_ptr()), st))
        got_ix = ix.cpu().numpy().view(np.uint64).reshape(W, G)
        got_dv = dv.cpu().numpy().view(np.uint64).reshape(W, 2 * G)
        assert np.array_equal(got_ix, ind[r0:r1]), f"indices [{r0}, {r1})"
        assert np.array_equal(got_dv, dat[r0:r1]), f"values [{r0}, {r1})"


@pytest.mark.parametrize("n,T,seed", [(15, 140, 1), (14, 300, 2), (15, 520, 3), (13, 64, 4),
                                      (14, 1100, 5)])
def test_mid_size_ops_bit_exact(n, T, seed):
    """Operators of 50..850 groups at 13..15 qubits (the single-operator build
    runs 129..1024 groups as two chunks on half the lanes per row)."""
    rng = np.random.default_rng(2000 + seed)
    labels = _random_labels(rng, n, T, prefixes=False)
    coeffs = rng.standard_normal(T) + 1j * rng.standard_normal(T)
    p = orc.Params(labels, coeffs)
    print(f"[gpu] mid-size op n={n} T={T} G={len(set(p.x.tolist()))}", flush=True)
    ip, ind, dat = orc.build_csr(p)
    shape, data, indices, indptr = _export(_op(labels, coeffs).to_matrix_mode("Hip"))
    assert indptr.tobytes() == ip.tobytes()
    assert indices.tobytes() == ind.tobytes()
    assert data.tobytes() == dat.tobytes()
