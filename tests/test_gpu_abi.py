"""The C ABI alone, as a binding without its own HIP layer uses it
(the Rust FFI block of the integration notes): qr_op_create / qr_op_info,
device buffers from qr_device_alloc, qr_csr_build and qr_csr_matvec on the
null stream, results copied back with qr_copy_to_host, checked against the
oracle (CSR bit-exact, y within 1e-12).  No torch memory is involved."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    yield


def test_c_abi_build_and_matvec(hamiltonians):
    from qrusty_amd import _lib as L
    lib = L.lib
    labels, coeffs = hamiltonians["rs_H6"]
    arr = (ctypes.c_char_p * len(labels))(*[s.encode() for s in labels])
    c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.complex128)).view(np.float64)
    op = ctypes.c_void_p()
    assert lib.qr_op_create(arr, c.ctypes.data_as(L.f64p), len(labels), L.QR_PHASE_ROWWISE,
                            ctypes.byref(op)) == 0
    nq, T, G, nnz = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    assert lib.qr_op_info(op, ctypes.byref(nq), ctypes.byref(T), ctypes.byref(G), ctypes.byref(nnz)) == 0
    N = 1 << nq.value
    bufs = {}
    for name, nbytes in [("ip", 8 * (N + 1)), ("ix", 8 * nnz.value), ("d", 16 * nnz.value),
                         ("x", 16 * N), ("y", 16 * N)]:
        p = ctypes.c_void_p()
        assert lib.qr_device_alloc(nbytes, ctypes.byref(p)) == 0 and p.value
        bufs[name] = p
    try:
        print(f"[gpu] c abi: N={N} nnz={nnz.value}", flush=True)
        assert lib.qr_csr_build(op, bufs["ip"], bufs["ix"], bufs["d"], None) == 0
        ip = np.empty(N + 1, dtype=np.uint64)
        ix = np.empty(nnz.value, dtype=np.uint64)
        dv = np.empty(nnz.value, dtype=np.complex128)
        for host, dev in [(ip, "ip"), (ix, "ix"), (dv, "d")]:
            assert lib.qr_copy_to_host(host.ctypes.data_as(ctypes.c_void_p), bufs[dev], host.nbytes,
                                       None) == 0
        p = orc.Params(labels, coeffs)
        rip, rix, rdv = orc.build_csr(p)
        assert np.array_equal(ip, rip) and np.array_equal(ix, rix)
        assert np.array_equal(dv.view(np.uint64), rdv.view(np.uint64))
        x = np.random.default_rng(5).standard_normal(N) + 1j * np.random.default_rng(6).standard_normal(N)
        assert lib.qr_copy_to_device(bufs["x"], x.ctypes.data_as(ctypes.c_void_p), x.nbytes, None) == 0
        assert lib.qr_csr_matvec(N, N, nnz.value, bufs["ip"], bufs["ix"], bufs["d"], bufs["x"],
                                 bufs["y"], None) == 0
        y = np.empty(N, dtype=np.complex128)
        assert lib.qr_copy_to_host(y.ctypes.data_as(ctypes.c_void_p), bufs["y"], y.nbytes, None) == 0
        ref = orc.spmat_dot_densevec(rip, rix, rdv, x)
        assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)
        # errors come back as codes with a message, not exceptions
        assert lib.qr_copy_to_host(None, bufs["y"], 16, None) == -1  # QR_ERR_ARG
        assert b"null" in lib.qr_last_error()
    finally:
        for p in bufs.values():
            assert lib.qr_device_free(p) == 0
        lib.qr_op_destroy(op)
