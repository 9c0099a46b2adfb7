"""GPU parity on the reference's giant Hamiltonians, the operators its timing
script builds (pyqrusty/tests/perf_giant.py:33-40, data from
pyqrusty/tests/H_fixtures.py as extracted into tests/golden):

  H10 20 qubits / 14251 terms / 2536 groups:  64 GB CSR, built whole
  H11 22 qubits / 20918 terms / 3792 groups: 382 GB CSR, larger than HBM
  H12 24 qubits /  4497 terms /  811 groups: 327 GB CSR, larger than HBM
  HJ  27 qubits /    55 terms /   28 groups:  90 GB CSR, built whole

CSR: the whole matrix where it fits (indptr closed form on every row, every
row's columns ascending and equal to {r ^ x_g}, sampled rows bit-exact against
the oracle's make_row, accel.rs:171-210, and the CSR matvec on sampled rows
within 1e-12); for H11 and H12 the row-streaming build (qr_csr_build_rows)
on three windows of rows (first, middle, last), every row's structure and
sampled rows' values bit-exact.  Matrix-free apply on all four: FP64 within
1e-12 of the oracle's rows of A x, mixed within 1e-6 of FP64.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

WHOLE = ["py_H10", "py_HJ"]
STREAMED = ["py_H11", "py_H12"]


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import qrusty_amd  # noqa: F401
    yield
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _op(labels, coeffs):
    import qrusty_amd as qa
    return qa.SparsePauliOp([qa.Pauli(l) for l in labels], coeffs)


def _abs_bound(p, x, rows):
    cabs = np.abs(p.cp.view(np.complex128))
    out = np.zeros(len(rows))
    for t in range(p.T):
        out += cabs[t] * np.abs(x[(rows ^ p.x[t]).astype(np.int64)])
    return out


def _check_rows(ind, r0, gx, step):
    """Rows r0.. of a (rows, G) device index block: ascending, and {r ^ x_g}."""
    n = ind.shape[0]
    G = ind.shape[1]
    for a in range(0, n, step):
        b = min(n, a + step)
        blk = ind[a:b]
        rows = torch.arange(r0 + a, r0 + b, device=ind.device, dtype=torch.int64)[:, None]
        assert bool((blk[:, 1:] > blk[:, :-1]).all()), f"rows {r0 + a}.. not ascending"
        assert torch.equal(torch.sort(blk ^ rows, dim=1).values, gx.expand(b - a, G)), \
            f"rows {r0 + a}.. columns"


def _sampled_values(p, ind_h, dat_h, rows, label):
    for i, r in enumerate(rows):
        cols, vals = orc.make_row(p, int(r))
        assert np.array_equal(ind_h[i], cols), f"{label} row {r} indices"
        assert np.array_equal(dat_h[i].view(np.float64), vals.view(np.float64)), f"{label} row {r} values"


@pytest.mark.parametrize("name", WHOLE)
def test_giant_csr_whole(hamiltonians, name):
    import qrusty_amd as qa
    labels, coeffs = hamiltonians[name]
    p = orc.Params(labels, coeffs)
    N = p.N
    gx_np = np.unique(p.x)
    G = len(gx_np)
    print(f"[gpu] giant CSR {name}: n={p.num_qubits} T={p.T} G={G} ({24 * N * G / 1e9:.1f} GB)",
          flush=True)
    op = _op(labels, coeffs)
    m = op.to_matrix_mode("RowwiseUnsafeChunked/1000")
    indptr, indices, data = m.device_arrays()
    dev = indices.device
    assert m.nnz() == N * G
    assert torch.equal(indptr, torch.arange(N + 1, device=dev, dtype=torch.int64) * G)
    gx = torch.from_numpy(gx_np.view(np.int64)).to(dev)
    ind = indices.view(N, G)
    _check_rows(ind, 0, gx, step=max(1, (1 << 25) // G))
    rows = np.unique(np.concatenate([[0, N - 1], np.random.default_rng(7).choice(N, 62, replace=False)]))
    sel = torch.from_numpy(rows.astype(np.int64)).to(dev)
    _sampled_values(p, ind[sel].cpu().numpy().view(np.uint64), data.view(N, G)[sel].cpu().numpy(),
                    rows, name)
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    x = torch.randn(N, dtype=torch.complex128, device="cuda", generator=g)
    x /= torch.linalg.vector_norm(x)
    y = qa.spmat_dot_densevec(m, x)
    xh = x.cpu().numpy()
    got = (y if isinstance(y, np.ndarray) else y.cpu().numpy())[rows]
    ref = orc.apply_rows(p, xh, rows.astype(np.uint64))[0]
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
    assert np.all(np.abs(got - ref) <= 1e-12 * _abs_bound(p, xh, rows.astype(np.uint64)) + 1e-300)
    del m, indptr, indices, data, ind, y
    torch.cuda.empty_cache()


@pytest.mark.parametrize("name", STREAMED)
def test_giant_csr_streamed_rows(hamiltonians, name):
    from qrusty_amd import _lib as L
    labels, coeffs = hamiltonians[name]
    p = orc.Params(labels, coeffs)
    N = p.N
    gx_np = np.unique(p.x)
    G = len(gx_np)
    print(f"[gpu] giant CSR {name} streamed: n={p.num_qubits} T={p.T} G={G} "
          f"({24 * N * G / 1e9:.1f} GB whole)", flush=True)
    op = _op(labels, coeffs)
    h = op._handle(L.QR_PHASE_ROWWISE)
    assert h.G == G
    W = 4096
    ind = torch.empty(W * G, dtype=torch.int64, device="cuda")
    dat = torch.empty(W * G, dtype=torch.complex128, device="cuda")
    gx = torch.from_numpy(gx_np.view(np.int64)).to(ind.device)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(13)
    for r0 in (0, N // 2 - W // 2 + 3, N - W):
        ind.fill_(-1)
        L.check(L.lib.qr_csr_build_rows(h.h, r0, r0 + W, ctypes.c_void_p(ind.data_ptr()),
                                        ctypes.c_void_p(dat.data_ptr()), s))
        torch.cuda.synchronize()
        iv = ind.view(W, G)
        _check_rows(iv, r0, gx, step=max(1, (1 << 24) // G))
        loc = np.unique(np.concatenate([[0, W - 1], rng.choice(W, 14, replace=False)]))
        sel = torch.from_numpy(loc.astype(np.int64)).to(ind.device)
        _sampled_values(p, iv[sel].cpu().numpy().view(np.uint64), dat.view(W, G)[sel].cpu().numpy(),
                        loc + r0, f"{name} window {r0}")
    del ind, dat
    torch.cuda.empty_cache()


@pytest.mark.parametrize("name", WHOLE + STREAMED)
def test_giant_apply(hamiltonians, name):
    labels, coeffs = hamiltonians[name]
    p = orc.Params(labels, coeffs)
    N = p.N
    print(f"[gpu] giant apply {name}: n={p.num_qubits} T={p.T}", flush=True)
    op = _op(labels, coeffs)
    g = torch.Generator(device="cuda")
    g.manual_seed(17)
    x = torch.randn(1, N, dtype=torch.complex128, device="cuda", generator=g)
    x /= torch.linalg.vector_norm(x)
    y64, e64 = op.apply_and_expectation(x, "fp64")
    y32, e32 = op.apply_and_expectation(x.to(torch.complex64), "mixed")
    torch.cuda.synchronize()
    rows = np.unique(np.concatenate([[0, N - 1], np.random.default_rng(19).choice(N, 46, replace=False)]))
    xh = x.cpu().numpy()
    sel = torch.from_numpy(rows.astype(np.int64)).to(y64.device)
    ref = orc.apply_rows(p, xh, rows.astype(np.uint64))[0]
    got = y64[0, sel].cpu().numpy()
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
    assert np.all(np.abs(got - ref) <= 1e-12 * _abs_bound(p, xh[0], rows.astype(np.uint64)) + 1e-300)
    ny = torch.linalg.vector_norm(y64).item()
    ry = torch.linalg.vector_norm(y32.to(torch.complex128) - y64).item() / ny
    e64, e32 = complex(np.ravel(e64)[0]), complex(np.ravel(e32)[0])
    vd = torch.sum(torch.conj(x) * y64).item()
    assert abs(e64 - vd) <= 1e-12 * ny
    de = abs(e32 - e64)
    print(f"[gpu] {name}: mixed rel err y {ry:.3e}, |dE| {de:.3e} (|E| {abs(e64):.3e}, "
          f"||Hx|| {ny:.3e})", flush=True)
    assert ry <= 1e-6
    assert de <= 1e-6 * ny
    del x, y64, y32
    torch.cuda.empty_cache()
