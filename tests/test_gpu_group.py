"""Term tables built on the device (qr_op_create_device) against the host path
(qr_op_create): label parse and make_params (qrusty/src/lib.rs:125-181,
accel.rs:141-157), the stable grouping by x mask (accel.rs:188-208) and the
rank lists must be bit-identical -- every table, the CSR rows built from
them, and the errors (code, message, which failing term is reported).
Operators: the reference's H11 (22 qubits, 20918 terms), a 200,000-term
3-local operator on 30 qubits (many terms per mask), 40-qubit random terms with
i/j/sign prefixes in both phase conventions, and small edge cases.
"""
import ctypes
import time

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


def _create(fn_name, labels, coeffs, conv):
    from qrusty_amd import _lib as L
    arr = (ctypes.c_char_p * len(labels))(*[s.encode() if s is not None else None for s in labels])
    c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.complex128)).view(np.float64)
    h = ctypes.c_void_p()
    args = [arr, c.ctypes.data_as(L.f64p), len(labels), conv]
    if fn_name == "qr_op_create_device":
        args.append(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    t0 = time.perf_counter()
    rc = getattr(L.lib, fn_name)(*args, ctypes.byref(h))
    dt = time.perf_counter() - t0
    msg = L.lib.qr_last_error().decode() if rc else ""
    return rc, msg, h, dt


def _tables(h):
    from qrusty_amd import _lib as L
    nq, T, G, nnz = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    L.check(L.lib.qr_op_info(h, ctypes.byref(nq), ctypes.byref(T), ctypes.byref(G), ctypes.byref(nnz)))
    T, G = T.value, G.value
    E = ctypes.c_uint64()
    L.check(L.lib.qr_op_tables(h, None, None, None, None, None, None, None, ctypes.byref(E)))
    out = {"nq": nq.value, "T": T, "G": G, "nnz": nnz.value,
           "gx": np.empty(G, np.uint64), "gterm": np.empty(G + 1, np.uint32),
           "tz": np.empty(T, np.uint64), "tc": np.empty(2 * T, np.float64),
           "gkind": np.empty(G, np.uint8), "rl_ptr": np.empty(G + 1, np.uint32),
           "rl": np.empty(2 * max(E.value, 1), np.int32)}
    L.check(L.lib.qr_op_tables(h, out["gx"].ctypes.data_as(L.u64p), out["gterm"].ctypes.data_as(L.u32p),
                               out["tz"].ctypes.data_as(L.u64p), out["tc"].ctypes.data_as(L.f64p),
                               out["gkind"].ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                               out["rl_ptr"].ctypes.data_as(L.u32p), out["rl"].ctypes.data_as(L.i32p),
                               None))
    x = np.empty(T, np.uint64)
    z = np.empty(T, np.uint64)
    cp = np.empty(2 * T, np.float64)
    L.check(L.lib.qr_op_terms(h, x.ctypes.data_as(L.u64p), z.ctypes.data_as(L.u64p),
                              cp.ctypes.data_as(L.f64p)))
    out.update({"x": x, "z": z, "cp": cp, "E": E.value, "rl": out["rl"][:2 * E.value]})
    return out


def _compare(labels, coeffs, conv, name):
    from qrusty_amd import _lib as L
    rc_h, _, hh, t_host = _create("qr_op_create", labels, coeffs, conv)
    rc_d, _, hd, t_dev = _create("qr_op_create_device", labels, coeffs, conv)
    _create("qr_op_create_device", labels, coeffs, conv)  # second run: no first-use costs
    _, _, hd2, t_dev2 = _create("qr_op_create_device", labels, coeffs, conv)
    assert rc_h == 0 and rc_d == 0
    try:
        a, b = _tables(hh), _tables(hd)
        for k in a:
            if isinstance(a[k], np.ndarray):
                assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), f"{name}: {k}"
            else:
                assert a[k] == b[k], f"{name}: {k}"
        # the row layouts (rank lists + sub-operator cuts) and CSR rows agree
        G, nq = a["G"], a["nq"]
        rng = np.random.default_rng(3)
        for r in [0, (1 << nq) - 1] + list(rng.integers(0, 1 << nq, 6)):
            c1, g1 = np.empty(G, np.uint64), np.empty(G, np.uint32)
            c2, g2 = np.empty(G, np.uint64), np.empty(G, np.uint32)
            L.check(L.lib.qr_op_row_layout(hh, int(r), c1.ctypes.data_as(L.u64p), g1.ctypes.data_as(L.u32p)))
            L.check(L.lib.qr_op_row_layout(hd, int(r), c2.ctypes.data_as(L.u64p), g2.ctypes.data_as(L.u32p)))
            assert np.array_equal(c1, c2) and np.array_equal(g1, g2)
        if nq <= 40:
            r0 = min(5, (1 << nq) - 1)
            W = min(64, (1 << nq) - r0)
            s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
            outs = []
            for h in (hh, hd):
                ix = torch.empty(W * G, dtype=torch.int64, device="cuda")
                dv = torch.empty(W * G, dtype=torch.complex128, device="cuda")
                L.check(L.lib.qr_csr_build_rows(h, r0, r0 + W, ctypes.c_void_p(ix.data_ptr()),
                                                ctypes.c_void_p(dv.data_ptr()), s))
                outs.append((ix, dv))
            torch.cuda.synchronize()
            assert torch.equal(outs[0][0], outs[1][0])
            assert torch.equal(outs[0][1].view(torch.int64), outs[1][1].view(torch.int64))
        print(f"[gpu] {name}: T={a['T']} G={G} E={a['E']} identical; create host {1e3 * t_host:.2f} ms, "
              f"device {1e3 * t_dev:.2f} ms (again {1e3 * t_dev2:.2f} ms)", flush=True)
    finally:
        for h in (hh, hd, hd2):
            L.lib.qr_op_destroy(h)


def _random_labels(rng, n, T, prefixes=False):
    codes = rng.integers(0, 4, size=(T, n))
    labels = ["".join("IXYZ"[c] for c in row) for row in codes]
    if prefixes:
        pre = ["", "-", "+", "i", "-j", "1", "-1i", "+j"]
        labels = [pre[k % len(pre)] + l for k, l in enumerate(labels)]
    return labels


def test_device_grouping_h11(hamiltonians):
    from qrusty_amd import _lib as L
    labels, coeffs = hamiltonians["py_H11"]
    print("[gpu] device grouping H11", flush=True)
    _compare(labels, coeffs, L.QR_PHASE_ROWWISE, "H11")


def test_device_grouping_large_k_local():
    from qrusty_amd import _lib as L
    from qrusty_amd import workloads
    labels, coeffs = workloads.k_local_op(30, 200_000, seed=91, max_weight=3)
    print("[gpu] device grouping 200k 3-local terms", flush=True)
    _compare(labels, coeffs, L.QR_PHASE_ROWWISE, "3-local 30q/200000")


@pytest.mark.parametrize("conv", [0, 1])
def test_device_grouping_prefixes_wide_masks(conv):
    rng = np.random.default_rng(92 + conv)
    labels = _random_labels(rng, 40, 30_000, prefixes=True)
    # repeated masks with other z and coefficient kinds
    labels += [l.replace("Y", "X") for l in labels[:3000]]
    coeffs = list(rng.standard_normal(len(labels)) + 1j * rng.standard_normal(len(labels)))
    for k in range(0, len(coeffs), 7):
        coeffs[k] = complex(coeffs[k].real, 0.0)
    for k in range(3, len(coeffs), 11):
        coeffs[k] = complex(0.0, coeffs[k].imag)
    coeffs[5] = complex(-0.0, 0.0)
    print(f"[gpu] device grouping 40q prefixes conv={conv}", flush=True)
    _compare(labels, coeffs, conv, f"40q prefixes conv={conv}")


def test_device_grouping_small_cases(hamiltonians):
    for name in ("py_H2", "py_H4", "rs_H6", "py_HJ"):
        labels, coeffs = hamiltonians[name]
        _compare(labels, coeffs, 0, name)
    _compare(["XYZ"], [1.5 - 2j], 0, "one term")
    _compare(["ZZ", "ZZ", "ZI"], [1.0, 2.0, 3.0], 1, "one mask")
    _compare(["X" * 63, "Y" * 63, "Z" * 63], [1.0, 1j, -1.0], 0, "63 qubits")


def test_device_grouping_errors():
    rng = np.random.default_rng(95)
    base = _random_labels(rng, 12, 2000)
    coeffs = [1.0] * 2000
    cases = {
        "malformed": base[:500] + ["IXQZ" + "I" * 8] + base[501:],
        "width then malformed": base[:300] + ["XY"] + base[301:500] + ["IXQZ" + "I" * 8] + base[501:],
        "malformed then width": base[:300] + ["-iXQ" + "I" * 8] + base[301:500] + ["XY"] + base[501:],
        "empty after prefix": base[:10] + ["-i"] + base[11:],
        "too long": base[:40] + ["X" * 64] + base[41:],
        "null": base[:700] + [None] + base[701:],
        "first term bad": ["Q"] + base[1:],
    }
    for name, labels in cases.items():
        rc_h, msg_h, hh, _ = _create("qr_op_create", labels, coeffs, 0)
        rc_d, msg_d, hd, _ = _create("qr_op_create_device", labels, coeffs, 0)
        assert rc_h != 0, name
        assert (rc_h, msg_h) == (rc_d, msg_d), name
