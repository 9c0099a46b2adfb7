"""qrusty_amd -- qrusty's Pauli-operator path on AMD MI355X (gfx950).

The Python surface keeps pyqrusty's names and behaviour for this path
(pyqrusty/src/lib.rs:262-433, 476-491; pyqrusty/python/pyqrusty/__init__.py):

    Pauli(label)                     .num_qubits() .label() repr/str .to_matrix()
    SparsePauliOp(paulis, coeffs)    len, [i], [a:b:c], .num_qubits(), repr,
                                     .to_matrix(), .to_matrix_mode(mode="")
    SpMat                            .shape() .nnz() .diagonal() .export()
                                     SpMat.new_unchecked(shape, data, indices, indptr)
    spmat_dot_densevec(spmat, x)     y = A x (complex128)
    csr_matrix(spmat)                scipy.sparse.csr_matrix from export()

and adds the matrix-free path:

    SparsePauliOp.apply(x, precision="fp64" | "mixed")
    SparsePauliOp.expectation(psi, precision="fp64" | "mixed")

Every compute call runs the HIP kernels of libqrhip.so through its C ABI
(include/qrhip.h); there is no CPU fallback.  Matrices live in device memory
(torch tensors are used only as device buffers and for the current stream).
"""
from __future__ import annotations

import ctypes
import decimal
import math

import numpy as np

from . import _lib as L
from ._lib import QrError, QrIOError

__all__ = ["Pauli", "SparsePauliOp", "SpMat", "spmat_dot_densevec", "csr_matrix", "axpby", "axpy",
           "ax", "precond", "precond2", "a_spmat_p_b_spmat", "Batch", "QrError",
           "QrIOError"]
__version__ = "0.1.0"

_PRECISIONS = {"fp64": L.QR_PREC_FP64, "mixed": L.QR_PREC_MIXED}


# ---------------------------------------------------------------------------
# device plumbing
# ---------------------------------------------------------------------------
def _torch():
    import torch
    return torch


def _require_gpu():
    torch = _torch()
    if not torch.cuda.is_available():
        raise QrError("qrusty_amd: no HIP device is available (the kernels have no CPU fallback)")
    return torch


def _stream_handle(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _rust_f64(v: float) -> str:
    """Rust's Display for f64 (shortest round-trip digits, never an exponent),
    used by util::complex64_to_string (qrusty/src/util.rs:140-142)."""
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    s = format(decimal.Decimal(repr(v)), "f")
    if "." in s:
        s = s.rstrip("0").rstrip(".")
    return s


# ---------------------------------------------------------------------------
# Pauli
# ---------------------------------------------------------------------------
class Pauli:
    """A Pauli string (pyqrusty/src/lib.rs:262-303; qrusty/src/lib.rs:117-267)."""

    __slots__ = ("_full", "_paulis", "_nq", "_base_phase", "_phase", "_x", "_z")

    def __init__(self, label: str):
        if not isinstance(label, str):
            raise TypeError("Pauli label must be a str")
        nq, bp, ph = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        x, z = ctypes.c_uint64(), ctypes.c_uint64()
        L.check(L.lib.qr_pauli_parse(label.encode(), ctypes.byref(nq), ctypes.byref(bp),
                                     ctypes.byref(ph), ctypes.byref(x), ctypes.byref(z)))
        self._full = label
        self._nq = nq.value
        self._paulis = label[len(label) - nq.value:]
        self._base_phase = bp.value
        self._phase = ph.value
        self._x = x.value
        self._z = z.value

    def num_qubits(self) -> int:
        return self._nq

    def label(self) -> str:
        """The I/X/Y/Z characters only, without a phase prefix (lib.rs:150-154)."""
        return self._paulis

    def __repr__(self) -> str:
        return f"Pauli('{self._paulis}')"

    def __str__(self) -> str:
        return self._paulis

    def to_matrix(self) -> "SpMat":
        """CSR of this Pauli string, i^base_phase times the Kronecker product
        (lib.rs:203-213), built by the HIP rowwise kernel."""
        return SparsePauliOp([self], [1.0 + 0.0j])._build(L.QR_PHASE_KRON)

    # internal accessors (masks as in lib.rs:161-181)
    @property
    def x_mask(self) -> int:
        return self._x

    @property
    def z_mask(self) -> int:
        return self._z

    @property
    def phase(self) -> int:
        return self._phase


# ---------------------------------------------------------------------------
# device operator handle
# ---------------------------------------------------------------------------
# Operators with at least this many terms build their term tables on the
# device (qr_op_create_device) when one is present; the tables are identical.
DEVICE_GROUPING_MIN_TERMS = 10_000


class _OpHandle:
    def __init__(self, labels, coeffs, convention):
        arr = (ctypes.c_char_p * len(labels))(*[s.encode() for s in labels])
        c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.complex128)).view(np.float64)
        h = ctypes.c_void_p()
        torch = _torch() if len(labels) >= DEVICE_GROUPING_MIN_TERMS else None
        if torch is not None and torch.cuda.is_available():
            L.check(L.lib.qr_op_create_device(arr, c.ctypes.data_as(L.f64p), len(labels), convention,
                                              _stream_handle(torch), ctypes.byref(h)))
        else:
            L.check(L.lib.qr_op_create(arr, c.ctypes.data_as(L.f64p), len(labels), convention,
                                       ctypes.byref(h)))
        self.h = h
        nq, T, G, nnz = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        L.check(L.lib.qr_op_info(h, ctypes.byref(nq), ctypes.byref(T), ctypes.byref(G),
                                 ctypes.byref(nnz)))
        self.num_qubits, self.T, self.G, self.nnz = nq.value, T.value, G.value, nnz.value

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and L is not None and L.lib is not None:
            L.lib.qr_op_destroy(h)
            self.h = None


# ---------------------------------------------------------------------------
# SparsePauliOp
# ---------------------------------------------------------------------------
class SparsePauliOp:
    """Sum of (Pauli, complex) terms (pyqrusty/src/lib.rs:305-433;
    qrusty/src/lib.rs:345-593)."""

    def __init__(self, paulis, coeffs):
        paulis = list(paulis)
        coeffs = [complex(c) for c in coeffs]
        if len(paulis) != len(coeffs):
            raise QrError("SparsePauliOp::new: paulis and coeffs must have same length")
        if len(paulis) == 0:
            raise QrError("SparsePauliOp::new: at least one pauli must be supplied")
        for p in paulis:
            if not isinstance(p, Pauli):
                raise TypeError("SparsePauliOp: paulis must be Pauli objects")
        nq = paulis[0].num_qubits()
        if any(p.num_qubits() != nq for p in paulis):
            raise QrError("SparsePauliOp::new: all supplied paulis must have the same #qubits")
        self._members = list(zip(paulis, coeffs))
        self._handles = {}

    # -- reference API ------------------------------------------------------
    def __len__(self) -> int:
        return len(self._members)

    def __getitem__(self, idx):
        if isinstance(idx, slice):
            return [m for m in self._members[idx]]
        if not isinstance(idx, int):
            raise TypeError("SparsePauliOp indices must be int or slice")
        if not (0 <= idx < len(self._members)):
            raise QrError(f"__getitem__ called on invalid index {idx}")
        return self._members[idx]

    def num_qubits(self) -> int:
        return self._members[0][0].num_qubits()

    def __repr__(self) -> str:
        labels = "','".join(p.label() for p, _ in self._members)
        coeffs = ", ".join(f"{_rust_f64(c.real)}+{_rust_f64(c.imag)}j" for _, c in self._members)
        return f"SparsePauliOp('{labels}', [{coeffs}])"

    __str__ = __repr__

    def __add__(self, other: "SparsePauliOp") -> "SparsePauliOp":
        """Concatenation of the term lists (qrusty/src/lib.rs:588-593)."""
        out = SparsePauliOp.__new__(SparsePauliOp)
        out._members = list(self._members) + list(other._members)
        out._handles = {}
        return out

    def to_matrix(self) -> "SpMat":
        return self.to_matrix_mode("")

    def to_matrix_mode(self, mode: str = "") -> "SpMat":
        """CSR of the operator (pyqrusty/src/lib.rs:406-415).

        Every reference mode string is accepted (lib.rs:293-331) and served by
        the HIP rowwise kernel; the mode only selects the reference's phase
        convention for i/j-prefixed labels ("", Binary, Reduce, Rayon use
        Pauli::to_matrix's i^base_phase; the others the rowwise (-i)^base_phase).
        """
        kind, step, conv = ctypes.c_int32(), ctypes.c_uint64(), ctypes.c_int32()
        L.check(L.lib.qr_mode_parse(mode.encode(), ctypes.byref(kind), ctypes.byref(step),
                                    ctypes.byref(conv)))
        return self._build(conv.value)

    # -- internals ----------------------------------------------------------
    def _handle(self, convention: int) -> _OpHandle:
        h = self._handles.get(convention)
        if h is None:
            labels = [p._full for p, _ in self._members]
            coeffs = [c for _, c in self._members]
            h = _OpHandle(labels, coeffs, convention)
            self._handles[convention] = h
        return h

    def _build(self, convention: int) -> "SpMat":
        torch = _require_gpu()
        h = self._handle(convention)
        N = 1 << h.num_qubits
        if h.nnz == 0:
            raise QrError("CSR too large: 2^n * G overflows 64 bits")
        dev = torch.device("cuda", torch.cuda.current_device())
        indptr = torch.empty(N + 1, dtype=torch.int64, device=dev)
        indices = torch.empty(h.nnz, dtype=torch.int64, device=dev)
        data = torch.empty(h.nnz, dtype=torch.complex128, device=dev)
        L.check(L.lib.qr_csr_build(h.h, _ptr(indptr), _ptr(indices), _ptr(data),
                                   _stream_handle(torch)))
        return SpMat._from_device((N, N), indptr, indices, data)

    # -- matrix-free path (no reference counterpart) --------------------------
    def apply(self, x, precision: str = "fp64", out=None):
        """y = H x without building the matrix.

        x: (N,) or (nvec, N) states -- numpy complex128, or torch device tensors
        (complex128 for "fp64", complex64 for "mixed").  Returns the same kind
        of object; a numpy input gives a numpy complex128 result.
        """
        return self._apply(x, precision, out=out, want_y=True, want_e=False)[0]

    def expectation(self, psi, precision: str = "fp64"):
        """<psi|H|psi> per state, reduced in FP64 on the device (y is not stored)."""
        return self._apply(psi, precision, out=None, want_y=False, want_e=True)[1]

    def apply_and_expectation(self, x, precision: str = "fp64", out=None):
        """(y = H x, <x|H|x>) from one pass."""
        return self._apply(x, precision, out=out, want_y=True, want_e=True)

    def _apply(self, x, precision, out, want_y, want_e):
        if precision not in _PRECISIONS:
            raise QrError(f"unknown precision {precision!r} (expected 'fp64' or 'mixed')")
        torch = _require_gpu()
        h = self._handle(L.QR_PHASE_ROWWISE)
        N = 1 << h.num_qubits
        prec = _PRECISIONS[precision]
        dt = torch.complex64 if prec == L.QR_PREC_MIXED else torch.complex128
        is_np = not isinstance(x, torch.Tensor)
        if is_np:
            xa = np.asarray(x)
            squeeze = xa.ndim == 1
            xt = torch.from_numpy(np.ascontiguousarray(xa.astype(np.complex128, copy=False)))
            xt = xt.to(device="cuda").to(dt)
        else:
            squeeze = x.dim() == 1
            if x.dtype != dt or not x.is_cuda:
                raise QrError(f"apply: expected a device tensor of dtype {dt}")
            xt = x.contiguous()
        if xt.dim() == 1:
            xt = xt.reshape(1, -1)
        if xt.dim() != 2 or xt.shape[1] != N:
            raise QrError(f"apply: state length must be 2^{h.num_qubits} = {N}")
        nvec = xt.shape[0]
        y = None
        if want_y:
            y = out if out is not None else torch.empty_like(xt)
            if not isinstance(y, torch.Tensor) or not y.is_cuda or y.device != xt.device:
                raise QrError("apply: out must be a device tensor on the state's device")
            if y.shape != xt.shape or y.dtype != dt or not y.is_contiguous():
                raise QrError("apply: out has the wrong shape, dtype or layout")
            nb = xt.numel() * xt.element_size()
            if y.data_ptr() < xt.data_ptr() + nb and xt.data_ptr() < y.data_ptr() + nb:
                raise QrError("apply: out overlaps the state (the apply is not in place)")
        e = torch.empty(2 * nvec, dtype=torch.float64, device=xt.device) if want_e else None
        L.check(L.lib.qr_pauli_apply(h.h, prec, nvec, _ptr(xt),
                                     _ptr(y) if y is not None else None,
                                     _ptr(e) if e is not None else None, _stream_handle(torch)))
        ev = None
        if e is not None:
            ev = e.cpu().numpy().view(np.complex128)
            if squeeze:
                ev = ev[0]
        if y is not None:
            if squeeze:
                y = y.reshape(-1)
            if is_np:
                y = y.to(torch.complex128).cpu().numpy()
        return y, ev


# ---------------------------------------------------------------------------
# SpMat
# ---------------------------------------------------------------------------
class SpMat:
    """CSR matrix with uint64 indptr/indices and complex128 data held in device
    memory (pyqrusty/src/lib.rs:33-260).  export() moves it out, after which
    the object is dead, like the reference's Box<Option<CsMatI>>."""

    def __init__(self):
        raise TypeError("SpMat objects come from to_matrix_mode() or SpMat.new_unchecked()")

    @classmethod
    def _from_device(cls, shape, indptr, indices, data) -> "SpMat":
        m = cls.__new__(cls)
        m._shape = (int(shape[0]), int(shape[1]))
        m._indptr, m._indices, m._data = indptr, indices, data
        return m

    @staticmethod
    def new_unchecked(shape, data, indices, indptr) -> "SpMat":
        """Wraps CSR arrays without validation (pyqrusty/src/lib.rs:103-116).
        Host arrays are copied to the device; torch device tensors (complex128
        data, 64-bit integer indices and indptr, contiguous) are adopted as
        they are, without a copy."""
        torch = _require_gpu()
        if all(isinstance(a, torch.Tensor) and a.is_cuda for a in (data, indices, indptr)):
            ok = (data.dtype == torch.complex128 and data.is_contiguous() and
                  all(a.dtype in (torch.int64, torch.uint64) and a.is_contiguous()
                      for a in (indices, indptr)))
            if not ok:
                raise QrError("SpMat.new_unchecked: device tensors must be contiguous complex128 "
                              "data and 64-bit integer indices and indptr")
            return SpMat._from_device(tuple(shape), indptr.view(torch.int64),
                                      indices.view(torch.int64), data)
        d = np.ascontiguousarray(np.asarray(data, dtype=np.complex128))
        i = np.ascontiguousarray(np.asarray(indices).astype(np.uint64)).view(np.int64)
        p = np.ascontiguousarray(np.asarray(indptr).astype(np.uint64)).view(np.int64)
        dev = torch.device("cuda", torch.cuda.current_device())
        return SpMat._from_device(tuple(shape), torch.from_numpy(p.copy()).to(dev),
                                  torch.from_numpy(i.copy()).to(dev),
                                  torch.from_numpy(d.copy()).to(dev))

    def _live(self, msg: str):
        if self._data is None:
            raise QrError(msg)

    def shape(self):
        self._live("SpMat.shape(): matrix is already dropped")
        return self._shape

    def nnz(self) -> int:
        self._live("cannot get NNZ of an exported sparse matrix")
        return int(self._data.numel())

    def diagonal(self, device: bool = False):
        """Dense main diagonal (pyqrusty/src/lib.rs:118-125): entry (r, r) of each
        row, 0 where absent, computed by qr_csr_diagonal on the device.  Returns
        numpy complex128 like the reference, or the device tensor with device=True."""
        self._live("SpMat.diagonal(): already-exported sparse matrix")
        torch = _require_gpu()
        out = torch.empty(min(self._shape), dtype=torch.complex128, device=self._data.device)
        L.check(L.lib.qr_csr_diagonal(self._shape[0], self._shape[1], self._data.numel(),
                                      _ptr(self._indptr), _ptr(self._indices), _ptr(self._data),
                                      _ptr(out), _stream_handle(torch)))
        return out if device else out.cpu().numpy()

    def export(self):
        """((rows, cols), data complex128, indices uint64, indptr uint64); the
        SpMat is dead afterwards (pyqrusty/src/lib.rs:190-214)."""
        self._live("cannot export from an already-exported sparse matrix")
        data = self._data.cpu().numpy()
        indices = self._indices.cpu().numpy().view(np.uint64)
        indptr = self._indptr.cpu().numpy().view(np.uint64)
        shape = self._shape
        self._indptr = self._indices = self._data = None
        return shape, data, indices, indptr

    def device_arrays(self):
        """(indptr, indices, data) device tensors (int64 views of uint64 values)."""
        self._live("cannot access an exported sparse matrix")
        return self._indptr, self._indices, self._data

    def __add__(self, other: "SpMat") -> "SpMat":
        """self + other (pyqrusty/src/lib.rs:150-152, sprs's `&a + &b`), run as
        1*self + 1*other: bit-identical for finite entries."""
        return self._binop(other, "add", 1.0, 1.0)

    def __sub__(self, other: "SpMat") -> "SpMat":
        """self - other (pyqrusty/src/lib.rs:154-156), run as 1*self + (-1)*other:
        bit-identical for finite entries."""
        return self._binop(other, "subtract", 1.0, -1.0)

    def _binop(self, other, name, a, b):
        """pyqrusty/src/lib.rs:40-52 (SpMat::binop), same messages."""
        if self._data is None or other._data is None:
            raise QrError(f"cannot {name} already-exported sparse matrices")
        if self._shape != other._shape:
            raise QrError(f"cannot {name} matrices with different shapes")
        return _spadd(a, self, b, other)

    def scale(self, factor) -> None:
        """In place: every stored value times factor (pyqrusty/src/lib.rs:158-163;
        num-complex v * factor), through qr_ax on the device."""
        self._live("cannot scale already-exported sparse matrix")
        torch = _require_gpu()
        out = torch.empty_like(self._data)
        L.check(L.lib.qr_ax(self._data.numel(), _scalar(factor), _ptr(self._data), _ptr(out),
                            _stream_handle(torch)))
        self._data = out

    def count_zeros(self, tolerance: float = 1e-7) -> int:
        """Entries with |v| <= tolerance (pyqrusty/src/lib.rs:170-175,
        qrusty/src/util.rs:144-152), counted on the device."""
        self._live("cannot count zeroes of an exported sparse matrix")
        torch = _require_gpu()
        n = ctypes.c_uint64()
        L.check(L.lib.qr_csr_count_zeros(self._data.numel(), _ptr(self._data), float(tolerance),
                                         ctypes.byref(n), _stream_handle(torch)))
        return int(n.value)

    def eliminate_zeros(self, tolerance: float = 1e-7) -> "SpMat":
        """A new SpMat without the entries with |v| <= tolerance
        (pyqrusty/src/lib.rs:177-183, qrusty/src/util.rs:154-171), compacted
        on the device.  Rows keep their column order (the reference's triplet
        rebuild sorts columns, which leaves sorted rows unchanged)."""
        self._live("cannot eliminate zeroes of an exported sparse matrix")
        torch = _require_gpu()
        nnz = self._data.numel()
        ipo = torch.empty_like(self._indptr)
        ixo = torch.empty(max(nnz, 1), dtype=torch.int64, device=self._data.device)
        dvo = torch.empty(max(nnz, 1), dtype=torch.complex128, device=self._data.device)
        kept = ctypes.c_uint64()
        L.check(L.lib.qr_csr_eliminate_zeros(self._shape[0], nnz, _ptr(self._indptr),
                                             _ptr(self._indices), _ptr(self._data),
                                             float(tolerance), _ptr(ipo), _ptr(ixo), _ptr(dvo),
                                             ctypes.byref(kept), _stream_handle(torch)))
        k = int(kept.value)
        return SpMat._from_device(self._shape, ipo, ixo[:k], dvo[:k])

    # -- files ----------------------------------------------------------------
    def matrixmarket_write(self, save_path: str) -> None:
        """MatrixMarket coordinate complex general file ("*.gz": gzip);
        pyqrusty/src/lib.rs:216-229 (sprs io, util.rs:172-199)."""
        self._live("cannot write an already-destroyed sparse matrix")
        self._write_mm(save_path, 0)

    def matrixmarket_write_symmetric(self, save_path: str, mode: str = "general") -> None:
        """The same with a symmetry mode ("general", "symmetric", "hermitian",
        "skew-symmetric"): the lower triangle is written (pyqrusty/src/lib.rs:
        88-99, 248-259).  The matrix is taken to have that symmetry."""
        if mode not in L.QR_MM_SYMMETRY:
            raise QrError(f"matrixmarket_write_symmetric: unrecognized mode {mode}")
        self._live("cannot write an already-destroyed sparse matrix")
        self._write_mm(save_path, L.QR_MM_SYMMETRY[mode])

    def _write_mm(self, path, sym):
        torch = _require_gpu()
        L.check(L.lib.qr_csr_write_mm(str(path).encode(), self._shape[0], self._shape[1],
                                      self._data.numel(), _ptr(self._indptr), _ptr(self._indices),
                                      _ptr(self._data), sym, _stream_handle(torch)))

    @staticmethod
    def matrixmarket_read(read_path: str) -> "SpMat":
        """Reads a MatrixMarket coordinate file (any symmetry, complex / real /
        integer field; "*.gz": gzip) into a device CSR with sorted rows and
        summed duplicates (pyqrusty/src/lib.rs:231-246)."""
        torch = _require_gpu()
        rd = ctypes.c_void_p()
        rows, cols, nnz = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        L.check(L.lib.qr_mm_read_open(str(read_path).encode(), ctypes.byref(rd), ctypes.byref(rows),
                                      ctypes.byref(cols), ctypes.byref(nnz), _stream_handle(torch)))
        try:
            dev = torch.device("cuda", torch.cuda.current_device())
            indptr = torch.empty(rows.value + 1, dtype=torch.int64, device=dev)
            indices = torch.empty(nnz.value, dtype=torch.int64, device=dev)
            data = torch.empty(nnz.value, dtype=torch.complex128, device=dev)
            L.check(L.lib.qr_mm_read_take(rd, _ptr(indptr), _ptr(indices), _ptr(data),
                                          _stream_handle(torch)))
        finally:
            L.lib.qr_mm_read_close(rd)
        return SpMat._from_device((rows.value, cols.value), indptr, indices, data)

    def raw_write(self, path: str) -> None:
        """qrusty's raw binary CSR format (qrusty/src/rawio.rs:128-147);
        "*.gz": gzip."""
        self._live("cannot write an already-destroyed sparse matrix")
        torch = _require_gpu()
        L.check(L.lib.qr_csr_write_raw(str(path).encode(), self._shape[0], self._shape[1],
                                       self._data.numel(), _ptr(self._indptr), _ptr(self._indices),
                                       _ptr(self._data), _stream_handle(torch)))

    @staticmethod
    def raw_read(path: str) -> "SpMat":
        """Reads qrusty's raw binary CSR format, either byte order
        (qrusty/src/rawio.rs:149-179)."""
        torch = _require_gpu()
        rows, cols, nnz = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        p = str(path).encode()
        L.check(L.lib.qr_csr_read_raw_info(p, ctypes.byref(rows), ctypes.byref(cols),
                                           ctypes.byref(nnz)))
        dev = torch.device("cuda", torch.cuda.current_device())
        indptr = torch.empty(rows.value + 1, dtype=torch.int64, device=dev)
        indices = torch.empty(nnz.value, dtype=torch.int64, device=dev)
        data = torch.empty(nnz.value, dtype=torch.complex128, device=dev)
        L.check(L.lib.qr_csr_read_raw(p, rows.value, nnz.value, _ptr(indptr), _ptr(indices),
                                      _ptr(data), _stream_handle(torch)))
        return SpMat._from_device((rows.value, cols.value), indptr, indices, data)

    def __copy__(self) -> "SpMat":
        if self._data is None:
            m = SpMat.__new__(SpMat)
            m._shape, m._indptr, m._indices, m._data = self._shape, None, None, None
            return m
        return SpMat._from_device(self._shape, self._indptr.clone(), self._indices.clone(),
                                  self._data.clone())

    def __repr__(self) -> str:
        if self._data is None:
            return "<already-dropped sparse matrix of type Complex64>"
        r, c = self._shape
        return (f"<{r}x{c} sparse matrix of type Complex64\n\twith {self.nnz()} stored "
                "elements in Compressed Sparse Row format>")

    __str__ = __repr__


def spmat_dot_densevec(spmat: SpMat, x):
    """y = spmat @ x on the device (pyqrusty/src/lib.rs:476-491; accel.rs:338-370).

    x: 1-D numpy complex128 (returns numpy) or a complex128 device tensor
    (returns a device tensor on the current stream)."""
    if spmat._data is None:
        raise QrError("cannot multiply with an exported sparse matrix")
    torch = _require_gpu()
    rows, cols = spmat._shape
    is_np = not isinstance(x, torch.Tensor)
    if is_np:
        xa = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
        if xa.ndim != 1:
            raise QrError("spmat_dot_densevec: x must be one-dimensional")
        xt = torch.from_numpy(xa).to(spmat._data.device)
    else:
        if x.dtype != torch.complex128 or x.dim() != 1 or not x.is_cuda:
            raise QrError("spmat_dot_densevec: x must be a 1-D complex128 device tensor")
        xt = x.contiguous()
    if xt.numel() != cols:
        raise QrError(f"spmat_dot_densevec: dimension mismatch ({xt.numel()} != {cols})")
    y = torch.empty(rows, dtype=torch.complex128, device=xt.device)
    L.check(L.lib.qr_csr_matvec(rows, cols, spmat._data.numel(), _ptr(spmat._indptr),
                                _ptr(spmat._indices), _ptr(spmat._data), _ptr(xt), _ptr(y),
                                _stream_handle(torch)))
    return y.cpu().numpy() if is_np else y


# ---------------------------------------------------------------------------
# eigensolver-step operations (pyqrusty/src/lib.rs:436-468, 519-585;
# qrusty/src/accel.rs:374-393)
# ---------------------------------------------------------------------------
def _vec_in(torch, x, name):
    """A 1-D complex128 device tensor for x, and whether x was numpy."""
    if isinstance(x, torch.Tensor):
        if x.dtype != torch.complex128 or x.dim() != 1 or not x.is_cuda:
            raise QrError(f"{name}: expected a 1-D complex128 device tensor")
        return x.contiguous(), False
    xa = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
    if xa.ndim != 1:
        raise QrError(f"{name}: expected a one-dimensional array")
    return torch.from_numpy(xa).to(device="cuda"), True


def _scalar(c):
    c = complex(c)
    return (ctypes.c_double * 2)(c.real, c.imag)


def _vec_out(torch, z, as_np):
    return z.cpu().numpy() if as_np else z


def axpby(a, x, b, y):
    """z = a x + b y (qrusty/src/accel.rs:374-379)."""
    torch = _require_gpu()
    xt, np_in = _vec_in(torch, x, "axpby")
    yt, _ = _vec_in(torch, y, "axpby")
    if xt.numel() != yt.numel():
        raise QrError(f"axpby: length mismatch ({xt.numel()} != {yt.numel()})")
    z = torch.empty_like(xt)
    L.check(L.lib.qr_axpby(xt.numel(), _scalar(a), _ptr(xt), _scalar(b), _ptr(yt), _ptr(z),
                           _stream_handle(torch)))
    return _vec_out(torch, z, np_in)


def axpy(a, x, y):
    """z = a x + y (qrusty/src/accel.rs:381-386)."""
    torch = _require_gpu()
    xt, np_in = _vec_in(torch, x, "axpy")
    yt, _ = _vec_in(torch, y, "axpy")
    if xt.numel() != yt.numel():
        raise QrError(f"axpy: length mismatch ({xt.numel()} != {yt.numel()})")
    z = torch.empty_like(xt)
    L.check(L.lib.qr_axpy(xt.numel(), _scalar(a), _ptr(xt), _ptr(yt), _ptr(z),
                          _stream_handle(torch)))
    return _vec_out(torch, z, np_in)


def ax(a, x):
    """z = a x (qrusty/src/accel.rs:388-393)."""
    torch = _require_gpu()
    xt, np_in = _vec_in(torch, x, "ax")
    z = torch.empty_like(xt)
    L.check(L.lib.qr_ax(xt.numel(), _scalar(a), _ptr(xt), _ptr(z), _stream_handle(torch)))
    return _vec_out(torch, z, np_in)


def precond(m: SpMat, dx, e, tol: float):
    """dx / reg(diag(m) - e, tol) (pyqrusty/src/lib.rs:441-454, 569-579)."""
    if m._data is None:
        raise QrError("cannot call precond with an exported sparse matrix")
    torch = _require_gpu()
    dxt, np_in = _vec_in(torch, dx, "precond")
    n = min(m._shape)
    if dxt.numel() != n:
        raise QrError(f"precond: dx has length {dxt.numel()}, the diagonal {n}")
    z = torch.empty_like(dxt)
    L.check(L.lib.qr_precond(m._shape[0], m._shape[1], m._data.numel(), _ptr(m._indptr),
                             _ptr(m._indices), _ptr(m._data), _ptr(dxt), _scalar(e), float(tol),
                             _ptr(z), _stream_handle(torch)))
    return _vec_out(torch, z, np_in)


def precond2(diag, dx, e, tol: float):
    """dx / reg(diag - e, tol) (pyqrusty/src/lib.rs:456-468, 581-591)."""
    torch = _require_gpu()
    dt, np_in = _vec_in(torch, diag, "precond2")
    dxt, _ = _vec_in(torch, dx, "precond2")
    if dt.numel() != dxt.numel():
        raise QrError(f"precond2: length mismatch ({dt.numel()} != {dxt.numel()})")
    z = torch.empty_like(dxt)
    L.check(L.lib.qr_precond2(dt.numel(), _ptr(dt), _ptr(dxt), _scalar(e), float(tol), _ptr(z),
                              _stream_handle(torch)))
    return _vec_out(torch, z, np_in)


def a_spmat_p_b_spmat(a, spmat1: SpMat, b, spmat2: SpMat) -> SpMat:
    """a*spmat1 + b*spmat2 on the device (pyqrusty/src/lib.rs:499-517); rows
    must hold ascending columns; exact zeros are dropped."""
    if spmat1._data is None:
        raise QrError("cannot a*m1+b*m2 with an exported sparse matrix (first argument)")
    if spmat2._data is None:
        raise QrError("cannot a*m1+b*m2 with an exported sparse matrix (second argument)")
    if spmat1._shape != spmat2._shape:
        raise QrError(f"cannot a*m1+b*m2 with matrices with differing shapes "
                      f"({spmat1._shape} != {spmat2._shape})")
    return _spadd(a, spmat1, b, spmat2)


def _spadd(a, spmat1: SpMat, b, spmat2: SpMat) -> SpMat:
    """Count pass, indptr scan and fill pass of qr_csr_add_count / _fill."""
    torch = _require_gpu()
    rows, cols = spmat1._shape
    dev = spmat1._data.device
    indptr = torch.empty(rows + 1, dtype=torch.int64, device=dev)
    nnz = ctypes.c_uint64()
    common = (rows, _scalar(a), spmat1._data.numel(), _ptr(spmat1._indptr), _ptr(spmat1._indices),
              _ptr(spmat1._data), _scalar(b), spmat2._data.numel(), _ptr(spmat2._indptr),
              _ptr(spmat2._indices), _ptr(spmat2._data))
    L.check(L.lib.qr_csr_add_count(*common, _ptr(indptr), ctypes.byref(nnz), _stream_handle(torch)))
    indices = torch.empty(nnz.value, dtype=torch.int64, device=dev)
    data = torch.empty(nnz.value, dtype=torch.complex128, device=dev)
    L.check(L.lib.qr_csr_add_fill(*common, _ptr(indptr), _ptr(indices), _ptr(data),
                                  _stream_handle(torch)))
    return SpMat._from_device((rows, cols), indptr, indices, data)


def csr_matrix(m: SpMat):
    """scipy.sparse.csr_matrix of an SpMat (pyqrusty/python/pyqrusty/__init__.py:14-17)."""
    (shape, data, indices, indptr) = m.export()
    from scipy.sparse import csr_matrix as _csr
    return _csr((data, indices, indptr), shape=shape, dtype=complex)


# ---------------------------------------------------------------------------
# batches (no counterpart in the reference, which handles one operator at a time)
# ---------------------------------------------------------------------------
class Batch:
    """Several operators built into CSR and applied to one state each, in one
    graph launch of the persistent batch kernels (qr_batch_run).

    Output buffers are allocated once and reused by every run(), so the
    recorded graph is replayed; CSR outputs are kept only when keep_csr
    (operators with more than 6144 distinct x masks always keep theirs).
    """

    def __init__(self, ops, keep_csr: bool = True):
        torch = _require_gpu()
        self._ops = list(ops)
        if not self._ops:
            raise QrError("empty batch")
        self._h = [op._handle(L.QR_PHASE_ROWWISE) for op in self._ops]
        n = len(self._ops)
        arr = (ctypes.c_void_p * n)(*[h.h.value for h in self._h])
        b = ctypes.c_void_p()
        L.check(L.lib.qr_batch_create(arr, n, ctypes.byref(b)))
        self._b = b
        dev = torch.device("cuda", torch.cuda.current_device())
        self.x, self.y = [], []
        self._csr = []
        for h in self._h:
            N = 1 << h.num_qubits
            self.x.append(torch.zeros(N, dtype=torch.complex128, device=dev))
            self.y.append(torch.zeros(N, dtype=torch.complex128, device=dev))
            if keep_csr or h.G > 6144:
                self._csr.append((torch.empty(N + 1, dtype=torch.int64, device=dev),
                                  torch.empty(h.nnz, dtype=torch.int64, device=dev),
                                  torch.empty(h.nnz, dtype=torch.complex128, device=dev)))
            else:
                self._csr.append(None)
        P = ctypes.c_void_p * n
        self._pp = P(*[c[0].data_ptr() if c else None for c in self._csr])
        self._pi = P(*[c[1].data_ptr() if c else None for c in self._csr])
        self._pd = P(*[c[2].data_ptr() if c else None for c in self._csr])
        self._px = P(*[t.data_ptr() for t in self.x])
        self._py = P(*[t.data_ptr() for t in self.y])

    def set_states(self, states) -> None:
        torch = _torch()
        for k, s in enumerate(states):
            t = s if isinstance(s, torch.Tensor) else torch.from_numpy(np.asarray(s, dtype=np.complex128))
            self.x[k].copy_(t.reshape(-1))

    def run(self, tuning: dict | None = None) -> None:
        """Build every CSR and y_k = A_k x_k (asynchronous on the current stream)."""
        torch = _torch()
        tu = None
        if tuning is not None:
            tu = L.QrTuning(int(tuning["tiles_per_item"]), int(tuning["blocks_per_cu"]),
                            int(tuning["order"]), int(tuning.get("split", 0)))
        L.check(L.lib.qr_batch_run(self._b, self._pp, self._pi, self._pd, self._px, self._py,
                                   ctypes.byref(tu) if tu is not None else None,
                                   _stream_handle(torch)))

    def tune(self) -> dict:
        """Time every tuning candidate on these buffers and keep the fastest."""
        torch = _torch()
        tu = L.QrTuning()
        L.check(L.lib.qr_batch_tune(self._b, self._pp, self._pi, self._pd, self._px, self._py,
                                    ctypes.byref(tu), _stream_handle(torch)))
        return {"tiles_per_item": tu.tiles_per_item, "blocks_per_cu": tu.blocks_per_cu,
                "order": tu.order, "split": tu.split}

    def csr(self, k: int):
        """(indptr, indices, data) device tensors of operator k, or None."""
        return self._csr[k]

    def __len__(self) -> int:
        return len(self._ops)

    def __del__(self):
        b = getattr(self, "_b", None)
        if b is not None and b.value and L is not None and L.lib is not None:
            L.lib.qr_batch_destroy(b)
            self._b = None
