"""CPU oracle for qrusty's Pauli-operator path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker (or the timed CPU
baseline), never as the thing measured or shipped.  The product path in
``qrusty_amd`` never imports it and fails loudly when its HIP library is
missing.

Two layers:

* ``liboracle.so`` (``qr_oracle.c``): a C restatement of the reference's
  rowwise build (qrusty/src/accel.rs:141-336) and CSR matvec
  (accel.rs:338-370), threaded like the reference (chunked dynamic scheduling).
* pure numpy/scipy pieces used to *pin* the restatement: a pure-Python label
  parser (qrusty/src/lib.rs:125-181) and a dense Kronecker build of a Pauli sum
  (lib.rs:93-114, 203-213, 401-412) that uses ``scipy.sparse.kron`` of the 2x2
  Pauli matrices -- a textbook routine independent of the x/z bit tricks.

* restatements of the CSR file formats (pure Python): ``rust_display_f64``
  (Rust's Display for f64), ``mm_text`` (qrusty/src/util.rs:172-199) and
  ``raw_bytes`` (qrusty/src/rawio.rs:128-147).

Parity status: every function is pinned by tests/test_oracle.py (known-answer
tests from lib.rs:608-813 and pyqrusty/tests/test_it.py, dense Kronecker
equality on the H2/H4/H6 fixtures as in pyqrusty/tests/test_H.py:32-48,
brute force on random small operators).  The bit order of the CSR matvec sum
(inside the absent third-party crate sprs) is "parity unpinned" at the bit
level and pinned to tolerance only.  The file formats: ``rust_display_f64``
is pinned by Rust's documented outputs, the reference's repr test ("1" for
1.0) and a round-trip plus shortest-digits check; ``mm_text`` by scipy's
mmread reading it back exactly; ``raw_bytes`` by an independent decode at
the documented offsets and the reference's rawio test matrix.  sprs's own
symmetric writer (write_matrix_market_sym; the crate is absent) is restated
as the MatrixMarket format stores a symmetric matrix, the lower triangle:
"parity unpinned" beyond that.
"""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

u64p = ctypes.POINTER(ctypes.c_uint64)
u32p = ctypes.POINTER(ctypes.c_uint32)
f64p = ctypes.POINTER(ctypes.c_double)


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_parse_label.argtypes = [ctypes.c_char_p, u32p, u32p, u64p, u64p, u32p]
        L.orc_parse_label.restype = ctypes.c_int
        L.orc_make_params.argtypes = [ctypes.c_uint64, u32p, f64p, f64p]
        L.orc_make_params.restype = None
        L.orc_make_row.argtypes = [ctypes.c_uint64, u64p, u64p, f64p, ctypes.c_uint64, u64p, f64p]
        L.orc_make_row.restype = ctypes.c_size_t
        L.orc_build_rows_chunked.argtypes = [ctypes.c_uint64, u64p, u64p, f64p, ctypes.c_uint64,
                                             ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                             ctypes.c_uint64, u64p, u64p, f64p]
        L.orc_build_rows_chunked.restype = ctypes.c_uint64
        L.orc_spmat_dot_densevec.argtypes = [ctypes.c_uint64, u64p, u64p, f64p, f64p, f64p, ctypes.c_int]
        L.orc_spmat_dot_densevec.restype = None
        L.orc_apply_rows.argtypes = [ctypes.c_uint64, u64p, u64p, f64p, ctypes.c_uint64, ctypes.c_uint32,
                                     f64p, u64p, ctypes.c_uint64, f64p, ctypes.c_int]
        L.orc_apply_rows.restype = None
        L.orc_apply_grouped.argtypes = [ctypes.c_uint64, u64p, u64p, u64p, f64p, ctypes.c_uint64,
                                        ctypes.c_uint32, f64p, f64p, ctypes.c_int]
        L.orc_apply_grouped.restype = None
        L.orc_vdot.argtypes = [ctypes.c_uint64, ctypes.c_uint32, f64p, f64p, f64p]
        L.orc_vdot.restype = None
        L.orc_csr_diagonal.argtypes = [ctypes.c_uint64, ctypes.c_uint64, u64p, u64p, f64p, f64p]
        L.orc_csr_diagonal.restype = None
        L.orc_axpby.argtypes = [ctypes.c_uint64, ctypes.c_int, f64p, f64p, f64p, f64p, f64p]
        L.orc_axpby.restype = None
        L.orc_precond2.argtypes = [ctypes.c_uint64, f64p, f64p, f64p, ctypes.c_double, f64p]
        L.orc_precond2.restype = None
        L.orc_csr_add.argtypes = [ctypes.c_uint64, f64p, u64p, u64p, f64p, f64p, u64p, u64p, f64p,
                                  u64p, u64p, f64p]
        L.orc_csr_add.restype = ctypes.c_uint64
        L.orc_count_zeros.argtypes = [ctypes.c_uint64, f64p, ctypes.c_double]
        L.orc_count_zeros.restype = ctypes.c_uint64
        L.orc_eliminate_zeros.argtypes = [ctypes.c_uint64, u64p, u64p, f64p, ctypes.c_double, u64p,
                                          u64p, f64p]
        L.orc_eliminate_zeros.restype = ctypes.c_uint64
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def default_threads() -> int:
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


# ---------------------------------------------------------------------------
# Label parsing
# ---------------------------------------------------------------------------
class OracleError(Exception):
    pass


def parse_label(label: str):
    """C restatement: returns (num_qubits, base_phase, x_mask, z_mask, phase)."""
    nq, bp, ph = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    x, z = ctypes.c_uint64(), ctypes.c_uint64()
    rc = lib().orc_parse_label(label.encode(), ctypes.byref(nq), ctypes.byref(bp),
                               ctypes.byref(x), ctypes.byref(z), ctypes.byref(ph))
    if rc != 0:
        raise OracleError("error: malformed label" if rc == -1 else "too many qubits")
    return nq.value, bp.value, x.value, z.value, ph.value


_LABEL_RE = re.compile(r"^([+-]?)1?([ij]?)([IXYZ]+)$")


def parse_label_py(label: str):
    """Pure-Python restatement with the reference's regex (lib.rs:127)."""
    m = _LABEL_RE.fullmatch(label)
    if m is None or label.endswith("\n"):
        raise OracleError("error: malformed label")
    sign, imag, paulis = m.group(1), m.group(2), m.group(3)
    base = (1 if imag else 0) + (2 if sign == "-" else 0)
    rev = paulis[::-1]
    x = sum(1 << q for q, c in enumerate(rev) if c in "XY")
    z = sum(1 << q for q, c in enumerate(rev) if c in "YZ")
    ny = sum(1 for c in rev if c == "Y")
    return len(rev), base, x, z, (base + ny) % 4


# ---------------------------------------------------------------------------
# make_params and the term table
# ---------------------------------------------------------------------------
class Params:
    """(x, z, c') per term in input order -- accel.rs:141-157."""

    def __init__(self, labels, coeffs):
        labels = list(labels)
        coeffs = np.asarray(coeffs, dtype=np.complex128).reshape(-1)
        if len(labels) != len(coeffs) or len(labels) == 0:
            raise OracleError("bad term list")
        parsed = [parse_label(l) for l in labels]
        nqs = {p[0] for p in parsed}
        if len(nqs) != 1:
            raise OracleError("mixed qubit counts")
        self.num_qubits = parsed[0][0]
        self.T = len(labels)
        self.x = np.array([p[2] for p in parsed], dtype=np.uint64)
        self.z = np.array([p[3] for p in parsed], dtype=np.uint64)
        self.base_phase = np.array([p[1] for p in parsed], dtype=np.uint32)
        self.phase = np.array([p[4] for p in parsed], dtype=np.uint32)
        cin = np.ascontiguousarray(coeffs.view(np.float64))
        cout = np.empty(2 * self.T, dtype=np.float64)
        lib().orc_make_params(self.T, _p(self.phase, u32p), _p(cin, f64p), _p(cout, f64p))
        self.cp = cout  # interleaved re, im

    @property
    def N(self) -> int:
        return 1 << self.num_qubits

    def groups(self):
        """Distinct x masks ascending, with each group's terms in input order."""
        order = np.argsort(self.x, kind="stable")
        xs = self.x[order]
        gx, first = np.unique(xs, return_index=True)
        gstart = np.append(first, self.T).astype(np.uint64)
        tz = np.ascontiguousarray(self.z[order])
        tc = np.ascontiguousarray(self.cp.reshape(-1, 2)[order].reshape(-1))
        return gx.astype(np.uint64), gstart, tz, tc


def make_row(params: Params, row: int):
    cols = np.empty(params.T, dtype=np.uint64)
    vals = np.empty(2 * params.T, dtype=np.float64)
    n = lib().orc_make_row(params.T, _p(params.x, u64p), _p(params.z, u64p), _p(params.cp, f64p),
                           row, _p(cols, u64p), _p(vals, f64p))
    return cols[:n].copy(), vals[: 2 * n].view(np.complex128).copy()


def build_csr(params: Params, row_begin: int = 0, row_end: int | None = None,
              step: int = 100, nthreads: int | None = None):
    """RowwiseUnsafeChunked/step over rows [row_begin, row_end).

    Returns (indptr, indices, data) with indptr counted from 0 for the slice.
    """
    if row_end is None:
        row_end = params.N
    nrows = row_end - row_begin
    G = len(np.unique(params.x))
    cap = nrows * G
    indptr = np.empty(nrows + 1, dtype=np.uint64)
    indices = np.empty(max(cap, 1), dtype=np.uint64)
    data = np.empty(2 * max(cap, 1), dtype=np.float64)
    nnz = lib().orc_build_rows_chunked(params.T, _p(params.x, u64p), _p(params.z, u64p),
                                       _p(params.cp, f64p), row_begin, row_end, step,
                                       nthreads or default_threads(), cap,
                                       _p(indptr, u64p), _p(indices, u64p), _p(data, f64p))
    if nnz == np.iinfo(np.uint64).max:
        raise OracleError("oracle build failed (capacity or arguments)")
    return indptr, indices[:nnz], data[: 2 * nnz].view(np.complex128)


def spmat_dot_densevec(indptr, indices, data, x, nthreads: int | None = None):
    nrows = len(indptr) - 1
    indptr = np.ascontiguousarray(indptr, dtype=np.uint64)
    indices = np.ascontiguousarray(indices, dtype=np.uint64)
    d = np.ascontiguousarray(np.asarray(data, dtype=np.complex128)).view(np.float64)
    xv = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).view(np.float64)
    y = np.empty(2 * nrows, dtype=np.float64)
    lib().orc_spmat_dot_densevec(nrows, _p(indptr, u64p), _p(indices, u64p), _p(d, f64p),
                                 _p(xv, f64p), _p(y, f64p), nthreads or default_threads())
    return y.view(np.complex128)


def apply_rows(params: Params, x, rows, nthreads: int | None = None):
    """(A x_v)[rows] for x of shape (nvec, N) via make_row + dot per row."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
    if x.ndim == 1:
        x = x[None, :]
    nvec, N = x.shape
    assert N == params.N
    rows = np.ascontiguousarray(rows, dtype=np.uint64)
    out = np.empty(2 * nvec * len(rows), dtype=np.float64)
    xv = x.view(np.float64).reshape(-1)
    lib().orc_apply_rows(params.T, _p(params.x, u64p), _p(params.z, u64p), _p(params.cp, f64p),
                         N, nvec, _p(xv, f64p), _p(rows, u64p), len(rows), _p(out, f64p),
                         nthreads or default_threads())
    return out.view(np.complex128).reshape(nvec, len(rows))


def apply_full(params: Params, x, nthreads: int | None = None):
    """y = A x for every row (FP64), same values as build_csr + spmat_dot_densevec."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
    if x.ndim == 1:
        x = x[None, :]
    nvec, N = x.shape
    assert N == params.N
    gx, gstart, tz, tc = params.groups()
    y = np.empty((nvec, N), dtype=np.complex128)
    lib().orc_apply_grouped(len(gx), _p(gx, u64p), _p(gstart, u64p), _p(tz, u64p), _p(tc, f64p),
                            N, nvec, _p(x.view(np.float64).reshape(-1), f64p),
                            _p(y.view(np.float64).reshape(-1), f64p), nthreads or default_threads())
    return y


def vdot(psi, y):
    """<psi_v|y_v> in FP64, row order."""
    psi = np.ascontiguousarray(np.asarray(psi, dtype=np.complex128))
    y = np.ascontiguousarray(np.asarray(y, dtype=np.complex128))
    if psi.ndim == 1:
        psi, y = psi[None, :], y[None, :]
    nvec, N = psi.shape
    out = np.empty(2 * nvec, dtype=np.float64)
    lib().orc_vdot(N, nvec, _p(psi.view(np.float64).reshape(-1), f64p),
                   _p(y.view(np.float64).reshape(-1), f64p), _p(out, f64p))
    return out.view(np.complex128)


# ---------------------------------------------------------------------------
# Dense Kronecker build (pins the restatement; independent of x/z bit tricks)
# ---------------------------------------------------------------------------
_PAULI_2x2 = {
    "I": np.array([[1, 0], [0, 1]], dtype=np.complex128),
    "X": np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    "Z": np.array([[1, 0], [0, -1]], dtype=np.complex128),
}


def pauli_kron_sparse(label: str):
    """Pauli string matrix by sparse Kronecker products, leftmost char most significant.

    lib.rs:203-213 folds kron(next_qubit, acc) starting from qubit 0, i.e. the
    matrix is kron(label[0], label[1], ..., label[-1]).  The base phase of a
    prefixed label is not applied here.
    """
    import scipy.sparse as sps
    m = _LABEL_RE.fullmatch(label)
    if m is None:
        raise OracleError("error: malformed label")
    paulis = m.group(3)
    acc = sps.csr_matrix(_PAULI_2x2[paulis[-1]])
    for c in reversed(paulis[:-1]):
        acc = sps.kron(sps.csr_matrix(_PAULI_2x2[c]), acc, format="csr")
    return acc, (1 if m.group(2) else 0) + (2 if m.group(1) == "-" else 0)


def dense_sum(labels, coeffs, convention: str = "rowwise"):
    """Dense sum_t c_t * phase(b_t) * P_t accumulated in term order.

    convention "rowwise": phase(b) = (-i)^b  (accel.rs:141-157, the rowwise modes)
    convention "kron":    phase(b) = i^b     (lib.rs:182-190, 211; to_matrix)
    For unprefixed labels (b = 0) both are 1.
    """
    coeffs = np.asarray(coeffs, dtype=np.complex128)
    first, base = pauli_kron_sparse(labels[0])
    N = first.shape[0]
    dense = np.zeros((N, N), dtype=np.complex128)
    for t, lab in enumerate(labels):
        P, b = pauli_kron_sparse(lab)
        ph = ((-1j) ** b) if convention == "rowwise" else ((1j) ** b)
        P = P.tocoo()
        vals = P.data * ph
        # each row of a Pauli string has exactly one nonzero: no duplicate (r, c)
        dense[P.row, P.col] = dense[P.row, P.col] + coeffs[t] * vals
    return dense


# ---------------------------------------------------------------------------
# eigensolver-step operations
# ---------------------------------------------------------------------------
def _c128(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128))


def csr_diagonal(shape, indptr, indices, data):
    """sprs CsMat::diag().to_dense() (pyqrusty/src/lib.rs:118-125)."""
    n = min(shape)
    out = np.zeros(n, dtype=np.complex128)
    ip = np.ascontiguousarray(indptr, dtype=np.uint64)
    ix = np.ascontiguousarray(indices, dtype=np.uint64)
    d = _c128(data)
    lib().orc_csr_diagonal(shape[0], shape[1], _p(ip, u64p), _p(ix, u64p),
                           _p(d.view(np.float64), f64p), _p(out.view(np.float64), f64p))
    return out


def axpby(a, x, b, y, mode: int = 0):
    """mode 0: a x + b y; 1: a x + y; 2: a x (qrusty/src/accel.rs:374-393)."""
    x = _c128(x)
    y = _c128(y) if y is not None else x
    z = np.empty_like(x)
    ab = np.array([complex(a).real, complex(a).imag, complex(b).real, complex(b).imag])
    lib().orc_axpby(len(x), mode, _p(ab[:2], f64p), _p(x.view(np.float64), f64p),
                    _p(np.ascontiguousarray(ab[2:]), f64p), _p(y.view(np.float64), f64p),
                    _p(z.view(np.float64), f64p))
    return z


def precond2(diag, dx, e, tol):
    """dx / reg(diag - e, tol) (pyqrusty/src/lib.rs:436-468)."""
    d = _c128(diag)
    dxa = _c128(dx)
    out = np.empty_like(dxa)
    ev = np.array([complex(e).real, complex(e).imag])
    lib().orc_precond2(len(d), _p(d.view(np.float64), f64p), _p(dxa.view(np.float64), f64p),
                       _p(ev, f64p), float(tol), _p(out.view(np.float64), f64p))
    return out


def csr_add(nrows, a, m1, b, m2):
    """a*A + b*B for CSR triples (indptr, indices, data) with ascending columns
    (pyqrusty/src/lib.rs:499-517).  Returns (indptr, indices, data)."""
    def arrs(m):
        ip, ix, d = m
        return (np.ascontiguousarray(ip, dtype=np.uint64), np.ascontiguousarray(ix, dtype=np.uint64),
                _c128(d))
    ip1, ix1, d1 = arrs(m1)
    ip2, ix2, d2 = arrs(m2)
    av = np.array([complex(a).real, complex(a).imag])
    bv = np.array([complex(b).real, complex(b).imag])
    ipo = np.empty(nrows + 1, dtype=np.uint64)
    args = (nrows, _p(av, f64p), _p(ip1, u64p), _p(ix1, u64p), _p(d1.view(np.float64), f64p),
            _p(bv, f64p), _p(ip2, u64p), _p(ix2, u64p), _p(d2.view(np.float64), f64p))
    nnz = lib().orc_csr_add(*args, _p(ipo, u64p), None, None)
    ixo = np.empty(max(nnz, 1), dtype=np.uint64)
    do = np.empty(max(nnz, 1), dtype=np.complex128)
    lib().orc_csr_add(*args, _p(ipo, u64p), _p(ixo, u64p), _p(do.view(np.float64), f64p))
    return ipo, ixo[:nnz], do[:nnz]


def count_zeros(data, tol):
    """Entries with hypot(re, im) <= tol (qrusty/src/util.rs:144-152)."""
    d = _c128(data)
    return int(lib().orc_count_zeros(len(d), _p(d.view(np.float64), f64p), tol))


def eliminate_zeros(nrows, m, tol):
    """CSR triple without the entries with hypot(re, im) <= tol
    (qrusty/src/util.rs:154-171)."""
    ip = np.ascontiguousarray(m[0], dtype=np.uint64)
    ix = np.ascontiguousarray(m[1], dtype=np.uint64)
    d = _c128(m[2])
    ipo = np.empty(nrows + 1, dtype=np.uint64)
    ixo = np.empty(max(len(d), 1), dtype=np.uint64)
    do = np.empty(max(len(d), 1), dtype=np.complex128)
    k = lib().orc_eliminate_zeros(nrows, _p(ip, u64p), _p(ix, u64p), _p(d.view(np.float64), f64p), tol,
                                  _p(ipo, u64p), _p(ixo, u64p), _p(do.view(np.float64), f64p))
    return ipo, ixo[:k], do[:k]


# ---------------------------------------------------------------------------
# CSR files
# ---------------------------------------------------------------------------
def rust_display_f64(v: float) -> str:
    """Rust's Display for f64: the shortest digits that read back to v
    (Python's repr has the same digits), written positionally, never with an
    exponent; "NaN", "inf", "-inf"; -0.0 gives "-0"."""
    if v != v:
        return "NaN"
    if v in (float("inf"), float("-inf")):
        return "inf" if v > 0 else "-inf"
    sign = "-" if str(v).startswith("-") else ""
    m, _, e = ("%r" % abs(v)).partition("e")
    digits = m.replace(".", "")
    point = (m.index(".") if "." in m else len(m)) + (int(e) if e else 0)
    digits = digits.lstrip("0")
    point -= len(m.replace(".", "")) - len(m.replace(".", "").lstrip("0"))
    digits = digits.rstrip("0")
    if not digits:
        return sign + "0"
    if point <= 0:
        return sign + "0." + "0" * (-point) + digits
    if point >= len(digits):
        return sign + digits + "0" * (point - len(digits))
    return sign + digits[:point] + "." + digits[point:]


_MM_SYM = {"general": 0, "symmetric": 1, "hermitian": 2, "skew-symmetric": 3}


def mm_text(shape, indptr, indices, data, symmetry="general") -> str:
    """MatrixMarket text as qrusty's util::write_matrix_market_for_scipy
    writes it (qrusty/src/util.rs:172-199): header, "% written by sprs",
    "rows cols nnz", then "row col re im" 1-based per entry in row order.  A
    symmetric mode names itself in the header and keeps the lower triangle
    (row >= col; row > col for skew-symmetric), as the MatrixMarket format
    stores symmetric matrices (sprs's write_matrix_market_sym; the crate is
    not in the reference tree)."""
    sym = _MM_SYM[symmetry]
    rows, cols = shape
    lines = []
    for r in range(rows):
        for e in range(int(indptr[r]), int(indptr[r + 1])):
            c = int(indices[e])
            if sym != 0 and (c > r or (sym == 3 and c == r)):
                continue
            v = complex(data[e])
            lines.append(f"{r + 1} {c + 1} {rust_display_f64(v.real)} {rust_display_f64(v.imag)}\n")
    return (f"%%MatrixMarket matrix coordinate complex {symmetry}\n% written by sprs\n"
            f"{rows} {cols} {len(lines)}\n" + "".join(lines))


def raw_bytes(shape, indptr, indices, data, byteorder="<") -> bytes:
    """qrusty::rawio::write (qrusty/src/rawio.rs:128-147): the bytes "MI" as
    the writer's u16 (so "IM" when read in the other order), then u64
    storage 0 (CSR), rows, cols, and the three arrays each after its u64
    length, all in the writer's byte order (`byteorder` "<" or ">")."""
    import struct
    mark = b"MI" if byteorder == "<" else b"IM"
    u = np.dtype(np.uint64).newbyteorder(byteorder)
    f = np.dtype(np.float64).newbyteorder(byteorder)
    d = np.ascontiguousarray(data, dtype=np.complex128).view(np.float64)
    out = [mark, struct.pack(byteorder + "4Q", 0, shape[0], shape[1], len(indptr)),
           np.asarray(indptr, dtype=np.uint64).astype(u).tobytes(),
           struct.pack(byteorder + "Q", len(indices)),
           np.asarray(indices, dtype=np.uint64).astype(u).tobytes(),
           struct.pack(byteorder + "Q", len(d) // 2), d.astype(f).tobytes()]
    return b"".join(out)
