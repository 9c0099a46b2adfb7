"""CPU tests of the product's host side: the C ABI library loads and exports
every declared symbol, and its host-only entry points (label parsing,
operator tables, row layout, mode strings) agree with the oracle.  No kernel
is launched here."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from oracle import oracle as orc

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(qr_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from qrusty_amd import _lib
    names = declared_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(_lib.lib, n), n
        assert n in _lib.SIGNATURES, f"{n} has no ctypes signature"
    assert set(_lib.SIGNATURES) == set(names)
    assert _lib.lib.qr_abi_version() == 1


def test_library_is_gfx950():
    # the offload bundle of the library built in this tree names its target
    from qrusty_amd import _lib
    with open(_lib.LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


# ---------------------------------------------------------------------------
def test_pauli_parse_matches_oracle_fuzz():
    from qrusty_amd import _lib as L
    rng = np.random.default_rng(11)
    alphabet = list("+-1ijIXYZQ")
    for _ in range(3000):
        s = "".join(rng.choice(alphabet, size=rng.integers(0, 8)))
        nq, bp, ph = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        x, z = ctypes.c_uint64(), ctypes.c_uint64()
        rc = L.lib.qr_pauli_parse(s.encode(), ctypes.byref(nq), ctypes.byref(bp), ctypes.byref(ph),
                                  ctypes.byref(x), ctypes.byref(z))
        try:
            ref = orc.parse_label(s)
        except orc.OracleError:
            ref = None
        if ref is None:
            assert rc == -2, s
            assert L.lib.qr_last_error() == b"error: malformed label"
        else:
            assert rc == 0 and (nq.value, bp.value, x.value, z.value, ph.value) == ref, s


def test_pauli_too_many_qubits():
    from qrusty_amd import Pauli, QrError
    Pauli("X" * 63)
    with pytest.raises(QrError):
        Pauli("X" * 64)


def test_op_terms_bit_exact_vs_make_params(hamiltonians):
    from qrusty_amd import SparsePauliOp, Pauli
    from qrusty_amd import _lib as L
    rng = np.random.default_rng(2)
    cases = [hamiltonians["py_H4"], hamiltonians["rs_H6"],
             (["iXY", "-jZZ", "-1YI", "+1iXX", "YY"],
              list(rng.standard_normal(5) + 1j * rng.standard_normal(5)))]
    for labels, coeffs in cases:
        op = SparsePauliOp([Pauli(l) for l in labels], coeffs)
        h = op._handle(L.QR_PHASE_ROWWISE)
        T = len(labels)
        x = np.empty(T, dtype=np.uint64)
        z = np.empty(T, dtype=np.uint64)
        c = np.empty(2 * T, dtype=np.float64)
        L.check(L.lib.qr_op_terms(h.h, x.ctypes.data_as(L.u64p), z.ctypes.data_as(L.u64p),
                                  c.ctypes.data_as(L.f64p)))
        p = orc.Params(labels, coeffs)
        assert np.array_equal(x, p.x) and np.array_equal(z, p.z)
        assert c.tobytes() == p.cp.tobytes()
        assert h.G == len(np.unique(p.x)) and h.nnz == p.N * h.G


def test_kron_convention_terms():
    from qrusty_amd import SparsePauliOp, Pauli
    from qrusty_amd import _lib as L
    labels = ["iX", "-jYZ", "-Y", "jI"]
    coeffs = [1.0, 2.0, 0.5 + 0.5j, -1.0]
    for lab, cf in zip(labels, coeffs):
        op = SparsePauliOp([Pauli(lab)], [cf])
        h = op._handle(L.QR_PHASE_KRON)
        c = np.empty(2, dtype=np.float64)
        L.check(L.lib.qr_op_terms(h.h, None, None, c.ctypes.data_as(L.f64p)))
        _, bp, _, _, _ = orc.parse_label(lab)
        ny = lab.count("Y")
        assert c.view(np.complex128)[0] == cf * (1j ** bp) * ((-1j) ** ny)


@pytest.mark.parametrize("name", ["py_H2", "py_H4", "py_H6", "py_H8", "py_H12"])
def test_row_layout_matches_make_row(hamiltonians, name):
    """The rank-list tables reproduce the stable-sorted column order of
    make_row (accel.rs:171-210) for random rows (no GPU involved)."""
    from qrusty_amd import SparsePauliOp, Pauli
    from qrusty_amd import _lib as L
    labels, coeffs = hamiltonians[name]
    op = SparsePauliOp([Pauli(l) for l in labels], coeffs)
    h = op._handle(L.QR_PHASE_ROWWISE)
    p = orc.Params(labels, coeffs)
    gx = np.unique(p.x)
    rng = np.random.default_rng(4)
    rows = list(rng.integers(0, p.N, size=40)) + [0, p.N - 1]
    for r in rows:
        cols = np.empty(h.G, dtype=np.uint64)
        grp = np.empty(h.G, dtype=np.uint32)
        L.check(L.lib.qr_op_row_layout(h.h, int(r), cols.ctypes.data_as(L.u64p),
                                       grp.ctypes.data_as(L.u32p)))
        ref_cols, _ = orc.make_row(p, int(r))
        assert np.array_equal(cols, ref_cols)
        assert np.array_equal(np.uint64(r) ^ gx[grp], cols)


def test_row_layout_degenerate_tries():
    """Chains (masks 1,2,4,...), full cubes and single groups."""
    from qrusty_amd import SparsePauliOp, Pauli
    from qrusty_amd import _lib as L
    n = 10
    fam = {
        "chain": ["I" * (n - 1 - q) + "X" + "I" * q for q in range(n)],
        "cube": ["".join("X" if (m >> (n - 1 - q)) & 1 else "I" for q in range(n)) for m in range(64)],
        "single": ["ZZZZZZZZZZ", "IIIIIIIIIZ"],
        "top": ["X" + "I" * (n - 1), "I" * n],
    }
    for name, labels in fam.items():
        op = SparsePauliOp([Pauli(l) for l in labels], [1.0] * len(labels))
        h = op._handle(L.QR_PHASE_ROWWISE)
        p = orc.Params(labels, [1.0] * len(labels))
        for r in range(0, 1 << n, 37):
            cols = np.empty(h.G, dtype=np.uint64)
            L.check(L.lib.qr_op_row_layout(h.h, r, cols.ctypes.data_as(L.u64p), None))
            assert np.array_equal(cols, orc.make_row(p, r)[0]), (name, r)


# ---------------------------------------------------------------------------
# modes (lib.rs:293-331, pyqrusty/src/lib.rs:406-415)
# ---------------------------------------------------------------------------
def _mode(s):
    from qrusty_amd import _lib as L
    k, st, cv = ctypes.c_int32(), ctypes.c_uint64(), ctypes.c_int32()
    rc = L.lib.qr_mode_parse(s.encode(), ctypes.byref(k), ctypes.byref(st), ctypes.byref(cv))
    return rc, k.value, st.value, cv.value


def test_mode_parse():
    assert _mode("")[1:] == (0, 0, 1)
    assert _mode("Binary")[1] == 1 and _mode("Binary")[3] == 1
    assert _mode("Accel")[1:] == (2, 0, 0)
    assert _mode("Rowwise")[1] == 3
    assert _mode("RowwiseUnsafe")[1] == 4
    assert _mode("RowwiseUnsafeChunked/100")[1:] == (5, 100, 0)
    assert _mode("xxRowwiseUnsafeChunked/12yy")[1:] == (5, 12, 0)   # unanchored is_match
    assert _mode("Reduce")[1:3] == (6, 0)
    assert _mode("Rayon")[1:] == (7, 0, 1)
    assert _mode("RayonChunked/250")[1:] == (8, 250, 0)
    assert _mode("Hip")[1:] == (9, 0, 0)
    for bad in ["foo", "rowwise", "RowwiseUnsafeChunked", "RowwiseUnsafeChunked/", "Accel ",
                "RowwiseUnsafeChunked/0"]:
        assert _mode(bad)[0] == -4, bad


def test_bad_mode_raises_like_reference(hamiltonians):
    from qrusty_amd import SparsePauliOp, Pauli, QrError
    labels, coeffs = hamiltonians["py_H2"]
    op = SparsePauliOp([Pauli(l) for l in labels], coeffs)
    with pytest.raises(Exception, match="to_matrix_mode: unrecognized mode foo"):
        op.to_matrix_mode("foo")


# ---------------------------------------------------------------------------
# Python surface that needs no GPU (pyqrusty/tests/test_it.py:28-44, 211-216)
# ---------------------------------------------------------------------------
def test_paulis_api():
    from qrusty_amd import Pauli
    p = Pauli("IXYZ")
    assert p.num_qubits() == 4
    assert p.label() == "IXYZ"
    assert repr(p) == "Pauli('IXYZ')"
    assert str(Pauli("-1jXY")) == "XY"
    with pytest.raises(Exception):
        Pauli("W")


def test_spop_api():
    from qrusty_amd import Pauli, SparsePauliOp
    with pytest.raises(Exception):
        SparsePauliOp([Pauli("I"), Pauli("IXYZ")], [1.0 + 0.0j, 1.0 + 0.0j])
    with pytest.raises(Exception, match="same length"):
        SparsePauliOp([Pauli("I")], [1.0, 2.0])
    with pytest.raises(Exception, match="at least one"):
        SparsePauliOp([], [])
    spop = SparsePauliOp([Pauli("IIII"), Pauli("IXYZ")], [1.0 + 0.0j, 1.0 + 0.0j])
    assert repr(spop) == "SparsePauliOp('IIII','IXYZ', [1+0j, 1+0j])"
    assert len(spop) == 2 and spop.num_qubits() == 4
    s2 = SparsePauliOp([Pauli("X")], [-0.5 - 1e-8j])
    assert repr(s2) == "SparsePauliOp('X', [-0.5+-0.00000001j])"
    # `+` concatenates the term lists, a's terms first (qrusty/src/lib.rs:588-593)
    s3 = SparsePauliOp([Pauli("IXXI")], [2.0 + 0.0j]) + spop
    assert repr(s3) == "SparsePauliOp('IXXI','IIII','IXYZ', [2+0j, 1+0j, 1+0j])"
    assert len(s3) == 3 and len(spop) == 2


def test_slice_repr(hamiltonians):
    from qrusty_amd import Pauli, SparsePauliOp
    labels, coeffs = hamiltonians["py_H2"]
    H2 = SparsePauliOp([Pauli(l) for l in labels], coeffs)
    assert repr(H2[3]) == "(Pauli('IIXX'), (-0.00170526+0j))"
    assert repr(H2[3:8:2]) == ("[(Pauli('IIXX'), (-0.00170526+0j)), (Pauli('IZII'), "
                               "(0.18388659+0j)), (Pauli('XXII'), (-0.00170526+0j))]")
    with pytest.raises(Exception):
        H2[len(H2)]
    both = H2 + H2
    assert len(both) == 2 * len(H2)


def test_apply_rejects_aliased_y_before_any_device_call():
    # the overlap check runs before the library touches a device, so the
    # (never dereferenced) addresses below only exercise the host check
    from qrusty_amd import _lib as L
    labels = (ctypes.c_char_p * 2)(b"XZI", b"ZZY")
    coeffs = (ctypes.c_double * 4)(1.0, 0.0, 0.5, -0.25)
    op = ctypes.c_void_p()
    L.check(L.lib.qr_op_create(labels, coeffs, 2, L.QR_PHASE_ROWWISE, ctypes.byref(op)))
    try:
        base = 0x7F0000000000
        for prec, elem in ((L.QR_PREC_FP64, 16), (L.QR_PREC_MIXED, 8)):
            nbytes = 2 * 8 * elem  # 2 vectors of 2^3 rows
            for dy in (0, elem, nbytes - elem, -(nbytes - elem)):
                rc = L.lib.qr_pauli_apply(op, prec, 2, ctypes.c_void_p(base),
                                          ctypes.c_void_p(base + dy), None, None)
                assert rc == -1, (prec, dy)
                assert b"overlaps" in L.lib.qr_last_error()
    finally:
        L.lib.qr_op_destroy(op)
