"""GPU test of the allocation-free apply (qr_pauli_apply_noalloc), through the
C ABI: its y and <x|H|x> equal qr_pauli_apply's bit for bit (same kernels,
same deterministic reduction), on two calls in a row -- the second with a
smaller input -- in both precisions; sizes beyond the scratch array and
operators whose tables are not on the device are refused before any launch.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import qrusty_amd  # noqa: F401
    yield
    torch.cuda.synchronize()


def _handle(labels, coeffs):
    import qrusty_amd as qa
    from qrusty_amd import _lib as L
    op = qa.SparsePauliOp([qa.Pauli(l) for l in labels], coeffs)
    return op, op._handle(L.QR_PHASE_ROWWISE)


def _run(fn, h, prec, x, want_y):
    from qrusty_amd import _lib as L
    nvec = x.shape[0]
    y = torch.full_like(x, float("nan")) if want_y else None
    e = torch.full((2 * nvec,), float("nan"), dtype=torch.float64, device=x.device)
    s = torch.cuda.current_stream()
    L.check(fn(h.h, prec, nvec, ctypes.c_void_p(x.data_ptr()),
               ctypes.c_void_p(y.data_ptr()) if y is not None else None,
               ctypes.c_void_p(e.data_ptr()), ctypes.c_void_p(s.cuda_stream)))
    torch.cuda.synchronize()
    return (y.cpu().numpy() if y is not None else None), e.cpu().numpy()


@pytest.mark.parametrize("prec_name", ["fp64", "mixed"])
def test_noalloc_equals_apply_two_calls(hamiltonians, prec_name):
    from qrusty_amd import _lib as L
    from qrusty_amd import workloads
    prec = L.QR_PREC_FP64 if prec_name == "fp64" else L.QR_PREC_MIXED
    dt = torch.complex128 if prec_name == "fp64" else torch.complex64
    cases = [("py_H8", 3, 51, True), ("py_H6", 1, 52, True), ("py_H6", 2, 53, False)]
    for name, nvec, seed, want_y in cases:  # second and third calls: smaller inputs
        labels, coeffs = hamiltonians[name]
        op, h = _handle(labels, coeffs)
        x = torch.from_numpy(workloads.random_states(1 << h.num_qubits, nvec, seed)).to("cuda", dt)
        print(f"[gpu] noalloc {prec_name} {name} nvec={nvec} y={want_y}", flush=True)
        y_ref, e_ref = _run(L.lib.qr_pauli_apply, h, prec, x, want_y)  # uploads the tables
        y, e = _run(L.lib.qr_pauli_apply_noalloc, h, prec, x, want_y)
        assert np.array_equal(e, e_ref), (name, nvec)
        if want_y:
            assert np.array_equal(y, y_ref), (name, nvec)
        assert np.all(np.isfinite(e))


def test_noalloc_other_tile_size(monkeypatch, hamiltonians):
    """Complex64 applies that would leave CUs short of blocks take tiles of
    2^10 rows instead of 2^11, each size with tables of its own.  When only
    the other size's tables are on the device, the allocation-free call uses
    those, and equals the allocating call that uploaded them bit for bit."""
    from qrusty_amd import _lib as L
    from qrusty_amd import workloads
    labels, coeffs = hamiltonians["py_H8"]  # 16 qubits: 32 tiles of 2^11 rows
    op, h = _handle(labels, coeffs)
    x = torch.from_numpy(workloads.random_states(1 << h.num_qubits, 2, 54)).to("cuda", torch.complex64)
    print("[gpu] noalloc mixed, tables of the other tile size", flush=True)
    monkeypatch.setenv("QR_APPLY_SMALL", "0")
    y_ref, e_ref = _run(L.lib.qr_pauli_apply, h, L.QR_PREC_MIXED, x, True)  # 2^11-row tables
    monkeypatch.delenv("QR_APPLY_SMALL")
    y, e = _run(L.lib.qr_pauli_apply_noalloc, h, L.QR_PREC_MIXED, x, True)  # would take 2^10
    assert np.array_equal(y, y_ref) and np.array_equal(e, e_ref)
    y2_ref, e2_ref = _run(L.lib.qr_pauli_apply, h, L.QR_PREC_MIXED, x, True)  # uploads 2^10
    y2, e2 = _run(L.lib.qr_pauli_apply_noalloc, h, L.QR_PREC_MIXED, x, True)
    assert np.array_equal(y2, y2_ref) and np.array_equal(e2, e2_ref)
    assert np.linalg.norm(y2 - y) <= 1e-6 * np.linalg.norm(y)


def test_noalloc_refusals(hamiltonians):
    from qrusty_amd import _lib as L
    labels, coeffs = hamiltonians["py_H8"]  # 16 qubits: 64 FP64 tiles
    op, h = _handle(labels, coeffs)
    x = torch.zeros(1, 1 << 16, dtype=torch.complex128, device="cuda")
    e = torch.zeros(2, dtype=torch.float64, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    xp, ep = ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(e.data_ptr())
    # tables not uploaded yet on this device: refused, nothing launched
    rc = L.lib.qr_pauli_apply_noalloc(h.h, L.QR_PREC_FP64, 1, xp, None, ep, s)
    assert rc == L.QR_ERR_ARG and b"not on this device" in L.lib.qr_last_error()
    L.check(L.lib.qr_pauli_apply(h.h, L.QR_PREC_FP64, 1, xp, None, ep, s))
    # 4097 vectors x 64 tiles x 2 doubles > 2^19: refused before any launch
    # (the size check precedes the kernels, which would read 4097 vectors)
    rc = L.lib.qr_pauli_apply_noalloc(h.h, L.QR_PREC_FP64, 4097, xp, None, ep, s)
    assert rc == L.QR_ERR_SIZE and b"scratch array" in L.lib.qr_last_error()
    # the mixed tables of this operator were never uploaded
    rc = L.lib.qr_pauli_apply_noalloc(h.h, L.QR_PREC_MIXED, 1, xp, None, ep, s)
    assert rc == L.QR_ERR_ARG
    torch.cuda.synchronize()
