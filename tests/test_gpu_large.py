"""GPU parity at the full benchmark sizes.

* 20 qubits, 2000 random terms: the whole CSR (2^20 rows x 1996 groups,
  50 GB) is built; every row is checked on the device against the closed form
  (indptr[r] = r*G, row r holds the columns {r ^ x_g} in ascending order), and
  sampled rows bit-exactly against the oracle (indices and values), the matvec
  against the oracle rows within 1e-12.
* 24-qubit Heisenberg ring, 8 states: FP64 apply on sampled rows vs the
  oracle (1e-12), mixed vs FP64 (<= 1e-6), <psi|H|psi> vs an FP64 vdot.
* 28 qubits, 10000 k-local terms (4 GiB complex128 state): FP64 apply on
  sampled rows vs the oracle, mixed vs FP64 (<= 1e-6), FP64 <psi|H|psi> vs
  vdot(psi, H psi), mixed <psi|H|psi> within 1e-6 * ||H psi|| of FP64 (the
  expectation of a traceless operator in a random state is a cancelling sum,
  so its relative error is not meaningful; it is printed).
"""
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
    torch.cuda.empty_cache()


def _op(labels, coeffs):
    import qrusty_amd as qa
    return qa.SparsePauliOp([qa.Pauli(l) for l in labels], coeffs)


def _device_state(N, nvec, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.randn((nvec, N), dtype=torch.complex128, device="cuda", generator=g)
    return x / torch.linalg.vector_norm(x, dim=1, keepdim=True)


def _abs_bound(p, x, rows):
    cabs = np.abs(p.cp.view(np.complex128))
    out = np.zeros(len(rows))
    for t in range(p.T):
        out += cabs[t] * np.abs(x[(rows ^ p.x[t]).astype(np.int64)])
    return out


def test_csr_20q_2000_full_size():
    import qrusty_amd as qa
    from qrusty_amd import _lib as L
    from qrusty_amd import workloads
    labels, coeffs = workloads.random_op(20, 2000, seed=workloads.SEED_CFG2)
    print("[gpu] csr 20q/2000 full size", flush=True)
    op = _op(labels, coeffs)
    m = op.to_matrix_mode("Hip")
    indptr, indices, data = m.device_arrays()
    N = 1 << 20
    h = op._handle(L.QR_PHASE_ROWWISE)
    G = h.G
    assert m.nnz() == N * G
    dev = indices.device
    assert torch.equal(indptr, torch.arange(N + 1, device=dev, dtype=torch.int64) * G)
    gx = np.empty(h.T, dtype=np.uint64)
    L.check(L.lib.qr_op_terms(h.h, gx.ctypes.data_as(L.u64p), None, None))
    gx = torch.from_numpy(np.unique(gx).view(np.int64)).to(dev)
    ind = indices.view(N, G)
    step = 1 << 14
    for r0 in range(0, N, step):
        rows = torch.arange(r0, r0 + step, device=dev, dtype=torch.int64)[:, None]
        blk = ind[r0:r0 + step]
        assert bool((blk[:, 1:] > blk[:, :-1]).all()), f"rows {r0}.. not ascending"
        assert torch.equal(torch.sort(blk ^ rows, dim=1).values, gx.expand(step, G)), f"rows {r0}.."
    # values and matvec: sampled rows against the oracle
    p = orc.Params(labels, coeffs)
    rows = np.unique(np.concatenate([[0, N - 1], np.random.default_rng(5).choice(N, 126, replace=False)]))
    sel = torch.from_numpy(rows.astype(np.int64)).to(dev)
    gi = ind[sel].cpu().numpy().view(np.uint64)
    gd = data.view(N, G)[sel].cpu().numpy()
    for i, r in enumerate(rows):
        cols, vals = orc.make_row(p, int(r))
        assert np.array_equal(gi[i], cols)
        assert np.array_equal(gd[i].view(np.float64), vals.view(np.float64))
    x = _device_state(N, 1, seed=77)[0]
    y = qa.spmat_dot_densevec(m, x)
    xh = x.cpu().numpy()
    ref = orc.apply_rows(p, xh, rows.astype(np.uint64))[0]
    got = (y if isinstance(y, np.ndarray) else y.cpu().numpy())[rows]
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
    assert np.all(np.abs(got - ref) <= 1e-12 * _abs_bound(p, xh, rows.astype(np.uint64)) + 1e-300)
    del m, indptr, indices, data, ind
    torch.cuda.empty_cache()


def _apply_case(labels, coeffs, nq, nvec, seed, nrows):
    op = _op(labels, coeffs)
    p = orc.Params(labels, coeffs)
    N = 1 << nq
    x = _device_state(N, nvec, seed)
    y64, e64 = op.apply_and_expectation(x, "fp64")
    x32 = x.to(torch.complex64)
    y32, e32 = op.apply_and_expectation(x32, "mixed")
    torch.cuda.synchronize()
    rows = np.unique(np.concatenate([[0, N - 1], np.random.default_rng(seed).choice(N, nrows, replace=False)]))
    xh = x.cpu().numpy()
    sel = torch.from_numpy(rows.astype(np.int64)).to(y64.device)
    ref = orc.apply_rows(p, xh, rows.astype(np.uint64))
    got = y64[:, sel].cpu().numpy()
    for v in range(nvec):
        assert np.linalg.norm(got[v] - ref[v]) <= 1e-12 * np.linalg.norm(ref[v])
        assert np.all(np.abs(got[v] - ref[v]) <= 1e-12 * _abs_bound(p, xh[v], rows.astype(np.uint64)) + 1e-300)
    ny = torch.linalg.vector_norm(y64, dim=1)
    ry = (torch.linalg.vector_norm(y32.to(torch.complex128) - y64, dim=1) / ny).max().item()
    # FP64 expectation == vdot(x, y64) up to summation order
    e64c = e64.cpu().numpy() if not isinstance(e64, np.ndarray) else e64
    vd = torch.sum(torch.conj(x) * y64, dim=1).cpu().numpy()
    assert np.all(np.abs(e64c - vd) <= 1e-12 * ny.cpu().numpy())
    e32c = e32.cpu().numpy() if not isinstance(e32, np.ndarray) else e32
    de = np.abs(e32c - e64c)
    print(f"[gpu] n={nq} T={len(labels)}: mixed rel err y {ry:.3e}; <H> abs err {de.max():.3e} "
          f"(rel {np.max(de / np.abs(e64c)):.3e}, vs ||Hx|| {np.max(de / ny.cpu().numpy()):.3e})", flush=True)
    assert ry <= 1e-6
    assert np.all(de <= 1e-6 * ny.cpu().numpy())
    return ry


def test_heisenberg_24_full_size():
    from qrusty_amd import workloads
    labels, coeffs = workloads.heisenberg_ring(24, seed=workloads.SEED_CFG3)
    print("[gpu] heisenberg 24 full size", flush=True)
    _apply_case(labels, coeffs, 24, 8, seed=0x5152, nrows=254)
    torch.cuda.empty_cache()


def test_k_local_28_full_size():
    from qrusty_amd import workloads
    labels, coeffs = workloads.k_local_op(28, 10000, seed=workloads.SEED_CFG5)
    print("[gpu] k-local 28q/10000 full size", flush=True)
    _apply_case(labels, coeffs, 28, 1, seed=0x5153, nrows=62)
    torch.cuda.empty_cache()
