"""GPU parity of the matrix-free Pauli-sum apply and <x|H|x>.

FP64 is checked against the oracle (make_row + dot per row, which is the CSR
build followed by the CSR matvec): within 1e-12 relative (normwise, and
componentwise against sum_g |D_g(r)||x[r^x_g]|).  The mixed-precision path
(complex64 state, FP32 products with TwoSum-compensated sums, FP64 group
merge and FP64 <x|y>) is checked against FP64 at <= 1e-6 relative, and the
measured error is printed for every case.
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


def _op(labels, coeffs):
    import qrusty_amd as qa
    return qa.SparsePauliOp([qa.Pauli(l) for l in labels], coeffs)


def _abs_bound(p, x, rows):
    """sum_t |c'_t| |x[r ^ x_t]| for the given rows: the componentwise error
    scale of a sum of products whose terms may be merged in any order (the
    kernel merges D_g(r) per tile, not in the reference's term order)."""
    rows = np.asarray(rows, dtype=np.uint64)
    cabs = np.abs(p.cp.view(np.complex128))
    out = np.zeros((x.shape[0], len(rows)))
    for t in range(p.T):
        out += cabs[t] * np.abs(x[:, (rows ^ p.x[t]).astype(np.int64)])
    return out


def _states(N, nvec, seed):
    from qrusty_amd import workloads
    return workloads.random_states(N, nvec, seed)


CASES = [("py_H4", 3), ("py_H6", 2), ("rs_H6", 1), ("py_H8", 2)]


@pytest.mark.parametrize("name,nvec", CASES)
def test_fp64_apply_vs_oracle(hamiltonians, name, nvec):
    labels, coeffs = hamiltonians[name]
    print(f"[gpu] apply fp64 {name} nvec={nvec}", flush=True)
    p = orc.Params(labels, coeffs)
    x = _states(p.N, nvec, seed=31)
    op = _op(labels, coeffs)
    y, e = op.apply_and_expectation(x, "fp64")
    rows = np.arange(p.N, dtype=np.uint64) if p.N <= 4096 else \
        np.random.default_rng(0).choice(p.N, size=2048, replace=False).astype(np.uint64)
    ref = orc.apply_rows(p, x, rows)
    got = y[:, rows.astype(np.int64)]
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
    assert np.all(np.abs(got - ref) <= 1e-12 * _abs_bound(p, x, rows) + 1e-300)
    yfull = orc.apply_full(p, x)
    eref = orc.vdot(x, yfull)
    assert np.all(np.abs(e - eref) <= 1e-12 * np.abs(eref) + 1e-14)
    # expectation-only launch (y never written) gives the same numbers
    e2 = op.expectation(x, "fp64")
    assert np.array_equal(e, e2)


@pytest.mark.parametrize("name,nvec", CASES)
def test_mixed_vs_fp64(hamiltonians, name, nvec):
    labels, coeffs = hamiltonians[name]
    p = orc.Params(labels, coeffs)
    x = _states(p.N, nvec, seed=32)
    op = _op(labels, coeffs)
    y64, e64 = op.apply_and_expectation(x, "fp64")
    y32, e32 = op.apply_and_expectation(x, "mixed")
    ry = np.linalg.norm(y32 - y64) / np.linalg.norm(y64)
    re = np.max(np.abs(e32 - e64) / np.abs(e64))
    print(f"[gpu] mixed {name}: rel err y {ry:.3e}, <H> {re:.3e}", flush=True)
    assert ry <= 1e-6
    assert re <= 1e-6


@pytest.mark.parametrize("seed", range(6))
def test_random_ops_apply(seed):
    rng = np.random.default_rng(500 + seed)
    n = int(rng.integers(1, 13))
    T = int(rng.integers(1, 200))
    pre = ["", "-", "i", "-1j"]
    labels = [str(rng.choice(pre)) + "".join(rng.choice(list("IXYZ"), size=n)) for _ in range(T)]
    coeffs = rng.standard_normal(T) + 1j * rng.standard_normal(T)
    nvec = int(rng.integers(1, 11))
    print(f"[gpu] apply random seed={seed} n={n} T={T} nvec={nvec}", flush=True)
    p = orc.Params(labels, coeffs)
    x = _states(p.N, nvec, seed=seed)
    op = _op(labels, coeffs)
    y = op.apply(x, "fp64")
    if y.ndim == 1:
        y = y[None, :]
    rows = np.arange(p.N, dtype=np.uint64)
    ref = orc.apply_rows(p, x, rows)
    assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)
    assert np.all(np.abs(y - ref) <= 1e-12 * _abs_bound(p, x, rows) + 1e-300)
    ym = op.apply(x, "mixed")
    if ym.ndim == 1:
        ym = ym[None, :]
    assert np.linalg.norm(ym - ref) <= 1e-6 * np.linalg.norm(ref)


def test_apply_equals_csr_matvec(hamiltonians):
    import qrusty_amd as qa
    labels, coeffs = hamiltonians["py_H8"]
    op = _op(labels, coeffs)
    m = op.to_matrix_mode("Hip")
    x = _states(1 << 16, 1, seed=3)[0]
    y_csr = qa.spmat_dot_densevec(m, x)
    y_mf = op.apply(x, "fp64")
    assert np.linalg.norm(y_csr - y_mf) <= 1e-13 * np.linalg.norm(y_csr)


def test_device_tensors_and_errors(hamiltonians):
    import qrusty_amd as qa
    labels, coeffs = hamiltonians["py_H4"]
    op = _op(labels, coeffs)
    x = torch.from_numpy(_states(256, 2, seed=4)).cuda()
    y = op.apply(x, "fp64")
    assert y.is_cuda and y.dtype == torch.complex128 and y.shape == (2, 256)
    ym = op.apply(x.to(torch.complex64), "mixed")
    assert ym.dtype == torch.complex64
    with pytest.raises(qa.QrError):
        op.apply(x.to(torch.complex64), "fp64")
    with pytest.raises(qa.QrError):
        op.apply(x[:, :128].contiguous(), "fp64")
    with pytest.raises(qa.QrError):
        op.apply(x, "fp16")
    # out: a host tensor, or one overlapping the state, is refused before any launch
    with pytest.raises(qa.QrError):
        op.apply(x, "fp64", out=torch.empty(2, 256, dtype=torch.complex128))
    with pytest.raises(qa.QrError):
        op.apply(x, "fp64", out=x)
    big = torch.zeros(3, 256, dtype=torch.complex128, device="cuda")
    big[:2] = x
    with pytest.raises(qa.QrError):
        op.apply(big[:2], "fp64", out=big[1:])


def test_expectations_on_two_streams(hamiltonians):
    """Launches of one operator on two streams at once each take their own
    <x|y> workspace: both energies match the oracle."""
    labels, coeffs = hamiltonians["py_H8"]
    p = orc.Params(labels, coeffs)
    op = _op(labels, coeffs)
    xa = _states(p.N, 2, seed=41)
    xb = _states(p.N, 2, seed=42)
    ea_ref = orc.vdot(xa, orc.apply_full(p, xa))
    eb_ref = orc.vdot(xb, orc.apply_full(p, xb))
    import ctypes
    from qrusty_amd import _lib as L
    ta, tb = torch.from_numpy(xa).cuda(), torch.from_numpy(xb).cuda()
    op.expectation(ta, "fp64")  # tables uploaded before the concurrent part
    h = op._handle(L.QR_PHASE_ROWWISE)
    torch.cuda.synchronize()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    ea = [torch.empty(4, dtype=torch.float64, device="cuda") for _ in range(8)]
    eb = [torch.empty(4, dtype=torch.float64, device="cuda") for _ in range(8)]
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    for k in range(8):  # no synchronisation between the launches
        L.check(L.lib.qr_pauli_apply(h.h, L.QR_PREC_FP64, 2, P(ta), None, P(ea[k]),
                                     ctypes.c_void_p(sa.cuda_stream)))
        L.check(L.lib.qr_pauli_apply(h.h, L.QR_PREC_FP64, 2, P(tb), None, P(eb[k]),
                                     ctypes.c_void_p(sb.cuda_stream)))
    torch.cuda.synchronize()
    outs = {"a": [e.cpu().numpy().view(np.complex128) for e in ea],
            "b": [e.cpu().numpy().view(np.complex128) for e in eb]}
    for e in outs["a"]:
        assert np.all(np.abs(e - ea_ref) <= 1e-12 * np.abs(ea_ref) + 1e-14)
    for e in outs["b"]:
        assert np.all(np.abs(e - eb_ref) <= 1e-12 * np.abs(eb_ref) + 1e-14)


# The two kernel forms of the apply: "gather" (tile of x in LDS, partner tiles
# gathered from global memory per group) and "staged" (partner tiles loaded
# into LDS once per run of groups that share them).  QR_APPLY_FORM forces one;
# the staged form needs full tiles (n >= 13 mixed, n >= 12 FP64).  Both are
# checked against the oracle on operators whose masks reach well above the
# tile, with many groups per partner tile (k-local) and with few (random,
# Heisenberg ring).
STAGED_CASES = [("klocal", 16, 400, 71), ("klocal", 14, 3000, 72), ("random", 14, 150, 73),
                ("random", 13, 40, 74), ("heis", 15, 0, 75)]


@pytest.mark.parametrize("kind,n,T,seed", STAGED_CASES)
def test_apply_forms_vs_oracle(monkeypatch, kind, n, T, seed):
    from qrusty_amd import workloads
    if kind == "klocal":
        labels, coeffs = workloads.k_local_op(n, T, seed=seed)
    elif kind == "random":
        labels, coeffs = workloads.random_op(n, T, seed=seed)
    else:
        labels, coeffs = workloads.heisenberg_ring(n, seed=seed)
    p = orc.Params(labels, coeffs)
    nvec = 2
    x = _states(p.N, nvec, seed=seed)
    rows = np.random.default_rng(seed).choice(p.N, size=2048, replace=False).astype(np.uint64)
    ref = orc.apply_rows(p, x, rows)
    yfull = orc.apply_full(p, x)
    eref = orc.vdot(x, yfull)
    escale = np.linalg.norm(x, axis=1) * np.linalg.norm(yfull, axis=1)  # |<x|Hx>| <= this
    for form in ("gather", "staged"):
        monkeypatch.setenv("QR_APPLY_FORM", "1" if form == "gather" else "2")
        print(f"[gpu] apply form={form} {kind} n={n} T={len(labels)} seed={seed}", flush=True)
        op = _op(labels, coeffs)
        y, e = op.apply_and_expectation(x, "fp64")
        got = y[:, rows.astype(np.int64)]
        assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
        assert np.all(np.abs(got - ref) <= 1e-12 * _abs_bound(p, x, rows) + 1e-300)
        assert np.all(np.abs(e - eref) <= 1e-12 * np.abs(eref) + 1e-14)
        assert np.array_equal(op.expectation(x, "fp64"), e)  # y never written
        # complex64: default tiles (2^11 rows) and small ones (2^10)
        for small in ("0", "1"):
            monkeypatch.setenv("QR_APPLY_SMALL", small)
            ym, em = op.apply_and_expectation(x, "mixed")
            ry = np.linalg.norm(ym[:, rows.astype(np.int64)] - ref) / np.linalg.norm(ref)
            print(f"[gpu]   mixed (small tiles {small}) rel err y {ry:.2e}", flush=True)
            assert ry <= 1e-6
            assert np.all(np.abs(em - eref) <= 1e-6 * escale)
        monkeypatch.delenv("QR_APPLY_SMALL")


def test_mixed_tile_size_choice(monkeypatch, hamiltonians):
    """Complex64 applies take 2^10-row tiles when 2^11-row ones would give
    fewer than 4 tiles per CU: one 16-qubit state (32 tiles) takes the small
    ones, so its results equal QR_APPLY_SMALL=1 bit for bit; 300 states of it
    (9600 tiles, more than 4 per CU on any MI355X) take the default ones and
    equal QR_APPLY_SMALL=0.  The two sizes agree within rounding."""
    import torch as _t
    labels, coeffs = hamiltonians["py_H8"]  # 16 qubits
    op = _op(labels, coeffs)
    cus = _t.cuda.get_device_properties(0).multi_processor_count
    for nvec, forced in [(1, "1"), (300, "0")]:
        assert ((1 << 16) >> 11) * nvec < 4 * cus if forced == "1" else ((1 << 16) >> 11) * nvec >= 4 * cus
        x = _states(1 << 16, nvec, seed=61)
        print(f"[gpu] mixed tile choice nvec={nvec}", flush=True)
        monkeypatch.delenv("QR_APPLY_SMALL", raising=False)
        y, e = op.apply_and_expectation(x, "mixed")
        monkeypatch.setenv("QR_APPLY_SMALL", forced)
        yf, ef = op.apply_and_expectation(x, "mixed")
        monkeypatch.setenv("QR_APPLY_SMALL", "1" if forced == "0" else "0")
        yo, _ = op.apply_and_expectation(x, "mixed")
        assert np.array_equal(y, yf) and np.array_equal(e, ef)
        assert np.linalg.norm(yo - y) <= 1e-6 * np.linalg.norm(y)
    monkeypatch.delenv("QR_APPLY_SMALL")


@pytest.mark.parametrize("form", ["1", "2"])
def test_wide_mask_instances(monkeypatch, form):
    """The 64-bit-mask kernel instances (used from 33 qubits on) on a 14-qubit
    operator, forced with QR_APPLY_WIDE, in both forms, against the oracle."""
    from qrusty_amd import workloads
    labels, coeffs = workloads.k_local_op(14, 300, seed=81)
    p = orc.Params(labels, coeffs)
    x = _states(p.N, 2, seed=82)
    monkeypatch.setenv("QR_APPLY_WIDE", "1")
    monkeypatch.setenv("QR_APPLY_FORM", form)
    print(f"[gpu] apply wide masks form={form}", flush=True)
    op = _op(labels, coeffs)
    y, e = op.apply_and_expectation(x, "fp64")
    rows = np.arange(p.N, dtype=np.uint64)
    ref = orc.apply_rows(p, x, rows)
    assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)
    assert np.all(np.abs(y - ref) <= 1e-12 * _abs_bound(p, x, rows) + 1e-300)
    eref = orc.vdot(x, orc.apply_full(p, x))
    assert np.all(np.abs(e - eref) <= 1e-12 * np.abs(eref) + 1e-14)
    for small in ("0", "1"):
        monkeypatch.setenv("QR_APPLY_SMALL", small)
        ym = op.apply(x, "mixed")
        assert np.linalg.norm(ym - ref) <= 1e-6 * np.linalg.norm(ref)


def _label_from_masks(n, xm, zm):
    """Label whose qubit q (last character = qubit 0) has x bit q of xm and z
    bit q of zm: I, X, Z, or Y when both."""
    chars = []
    for q in reversed(range(n)):
        xb, zb = (xm >> q) & 1, (zm >> q) & 1
        chars.append("IXZY"[xb | (zb << 1)])
    return "".join(chars)


@pytest.mark.parametrize("small", ["0", "1"])
@pytest.mark.parametrize("n,seed", [(11, 91), (12, 92), (13, 93)])
def test_mixed_gather_row_pairs(monkeypatch, n, seed, small):
    """The complex64 gather form on full tiles takes rows in adjacent pairs,
    with the patterns' low z bits permuted to match.  Masks here set row bit 0
    (pair swap), the thread bits 1..8 and the row bits 9..10 on both sides of
    that split, with complex coefficients; every row is compared with the
    oracle, normwise and componentwise.  On tiles of 2^11 rows (QR_APPLY_SMALL
    = 0) and of 2^10 (= 1: bit 10 then reaches the partner tile)."""
    rng = np.random.default_rng(seed)
    picks = [0, 1, 2, 3, 0x100, 0x101, 0x200, 0x201, 0x300, 0x401, 0x600, 0x603, 0x7FF]
    hi = [1 << b for b in range(11, n)]
    labels, coeffs = [], []
    for _ in range(48):
        xm = int(rng.choice(picks)) | (int(rng.choice(hi)) if hi and rng.random() < 0.4 else 0)
        zm = int(rng.choice(picks)) ^ (int(rng.choice(hi)) if hi and rng.random() < 0.4 else 0)
        labels.append(_label_from_masks(n, xm & ((1 << n) - 1), zm & ((1 << n) - 1)))
        coeffs.append(complex(rng.normal(), rng.normal()))
    from qrusty_amd import workloads
    hl, hc = workloads.heisenberg_ring(n, seed=seed)
    labels += hl
    coeffs += list(hc)
    p = orc.Params(labels, coeffs)
    x = _states(p.N, 2, seed=seed)
    rows = np.arange(p.N, dtype=np.uint64)
    ref = orc.apply_rows(p, x, rows)
    eref = orc.vdot(x, ref)
    monkeypatch.setenv("QR_APPLY_FORM", "1")
    monkeypatch.setenv("QR_APPLY_SMALL", small)
    print(f"[gpu] apply mixed gather row pairs n={n} T={len(labels)} seed={seed} small={small}", flush=True)
    op = _op(labels, coeffs)
    ym, em = op.apply_and_expectation(x, "mixed")
    ry = np.linalg.norm(ym - ref) / np.linalg.norm(ref)
    print(f"[gpu]   mixed rel err y {ry:.2e}", flush=True)
    assert ry <= 1e-6
    assert np.all(np.abs(ym - ref) <= 1e-5 * _abs_bound(p, x, rows) + 1e-30)
    for v in range(2):
        scale = np.linalg.norm(x[v]) * np.linalg.norm(ref[v])
        assert abs(em[v] - eref[v]) <= 1e-6 * scale
    # the expectation-only launch (y not written) gives the same value
    assert np.array_equal(op.expectation(x, "mixed"), em)
