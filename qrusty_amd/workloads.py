"""Seeded synthetic inputs for the Pauli-operator path (labels, coefficients, states).

Every generator is deterministic for a given seed (numpy PCG64), so the GPU
run, the CPU oracle and the committed golden fixtures see identical inputs.
Labels follow qrusty's convention: qubit q is character ``label[n-1-q]``
(qrusty/src/lib.rs:144).
"""
from __future__ import annotations

import numpy as np

SEED_CFG1 = 0x51520001   # 12 qubits, 200 random terms (CPU-runnable case)
SEED_CFG2 = 0x51520002   # 20 qubits, 2000 random terms
SEED_CFG3 = 0x51520003   # 24-qubit Heisenberg ring, 8 states
SEED_CFG4 = 0x51520004   # 256 uneven operators
SEED_CFG5 = 0x51520005   # 28 qubits, 10000 k-local terms

_PAULI_CHARS = np.array(list("IXYZ"))


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def random_op(num_qubits: int, num_terms: int, seed: int):
    """Labels i.i.d. uniform over {I,X,Y,Z} per qubit; c = N(0,1) + i N(0,1)."""
    g = _rng(seed)
    codes = g.integers(0, 4, size=(num_terms, num_qubits))
    labels = ["".join(row) for row in _PAULI_CHARS[codes]]
    re = g.standard_normal(num_terms)
    im = g.standard_normal(num_terms)
    return labels, list(re + 1j * im)


def _label(n: int, ops: dict) -> str:
    """ops: {qubit: 'X'|'Y'|'Z'}; identity elsewhere."""
    chars = ["I"] * n
    for q, c in ops.items():
        chars[n - 1 - q] = c
    return "".join(chars)


def heisenberg_ring(num_qubits: int, seed: int):
    """X_iX_{i+1}, Y_iY_{i+1}, Z_iZ_{i+1} (i = 0..n-1, periodic) plus Z_i fields;
    real coefficients uniform in [-1, 1]."""
    g = _rng(seed)
    n = num_qubits
    labels = []
    for i in range(n):
        j = (i + 1) % n
        for p in "XYZ":
            labels.append(_label(n, {i: p, j: p}))
    for i in range(n):
        labels.append(_label(n, {i: "Z"}))
    c = g.uniform(-1.0, 1.0, size=len(labels))
    return labels, [complex(v, 0.0) for v in c]


def k_local_op(num_qubits: int, num_terms: int, seed: int, max_weight: int = 4):
    """Weight k ~ U{1..max_weight}, support uniform without replacement,
    factors uniform over {X,Y,Z}, real coefficients N(0,1)."""
    g = _rng(seed)
    labels = []
    for _ in range(num_terms):
        k = int(g.integers(1, max_weight + 1))
        support = g.choice(num_qubits, size=k, replace=False)
        facs = g.integers(0, 3, size=k)
        labels.append(_label(num_qubits, {int(q): "XYZ"[int(f)] for q, f in zip(support, facs)}))
    c = g.standard_normal(num_terms)
    return labels, [complex(v, 0.0) for v in c]


def uneven_batch(num_ops: int, seed: int, qmin: int = 8, qmax: int = 18,
                 tmin: int = 10, tmax: int = 5000):
    """Qubits ~ U{qmin..qmax}, terms = floor(exp(U[ln tmin, ln tmax])), labels and
    coefficients as random_op."""
    g = _rng(seed)
    ops = []
    for k in range(num_ops):
        nq = int(g.integers(qmin, qmax + 1))
        nt = int(np.floor(np.exp(g.uniform(np.log(tmin), np.log(tmax)))))
        ops.append(random_op(nq, max(nt, 1), seed=seed * 1000 + k))
    return ops


def random_states(dim: int, nvec: int, seed: int) -> np.ndarray:
    """nvec complex128 states with N(0,1) real/imag parts, each L2-normalised."""
    g = _rng(seed)
    x = g.standard_normal((nvec, dim)) + 1j * g.standard_normal((nvec, dim))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    return x
