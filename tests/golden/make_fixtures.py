"""Extract the reference's own test Hamiltonians as data and derive golden outputs.

Run where the qrusty source tree is mounted (default /root/reference); the
tests only read the fixtures it wrote, never this script.  The reference files are read
as TEXT and parsed with regular expressions -- nothing from them is imported
or executed:

* pyqrusty/tests/H_fixtures.py: ``NAME = build_spop([labels...], coeffs=[...])``
  (H2 4q/27, H4 8q/361, H6 12q/1819, H8 16q/5793, H10 20q/14251, H11 22q/20918,
  H12 24q/4497, HJ 27q/55; the timed operators of pyqrusty/tests/perf_giant.py:33-40)
* qrusty/src/fixtures.rs: ``TestCase`` H2/H4/H6 with full-precision coefficient
  strings (lib.rs tests h2/h4/h6 use these).

Outputs (tests/golden/):
* ``hamiltonians.json``  -- {name: {"labels": [...], "coeffs": [[re, im], ...]}}
* ``csr_digests.json``   -- per fixture with <= 16 qubits: G, nnz and SHA-256
  digests of indptr/indices/data of the rowwise CSR, computed by the oracle
  after the oracle has been checked against a dense Kronecker build
  (n <= 12, exact equality as in pyqrusty/tests/test_H.py:32-48).
* ``config1.json`` / ``config1_y.npy`` -- the CPU-runnable workload (12 qubits,
  200 random terms, seed fixed here) with its CSR digests and y = A x.
* ``precond_h2.json`` -- dx, e, tol and the printed result vector of the
  precond tests in pyqrusty/tests/test_it.py:276-300 (inputs and a 9-digit
  known answer for H2).
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from oracle import oracle as orc  # noqa: E402
from qrusty_amd import workloads  # noqa: E402

PY_FIXTURES = ["H2", "H4", "H6", "H8", "H10", "H11", "H12", "HJ"]
RS_FIXTURES = ["H2", "H4", "H6"]


def parse_py_fixtures(text: str):
    out = {}
    for m in re.finditer(r"^(\w+) = build_spop\(\s*\[(.*?)\],\s*coeffs=\[(.*?)\],\s*\)", text,
                         re.S | re.M):
        name, lab_txt, coef_txt = m.group(1), m.group(2), m.group(3)
        labels = re.findall(r'"([^"]*)"', lab_txt)
        coeffs = [complex(c.replace(" ", "")) for c in coef_txt.split(",") if c.strip()]
        if len(labels) != len(coeffs):
            raise SystemExit(f"{name}: {len(labels)} labels vs {len(coeffs)} coeffs")
        out[name] = (labels, coeffs)
    return out


def parse_rs_fixtures(text: str):
    out = {}
    for m in re.finditer(r"pub static ref (\w+) : TestCase = \{(.*?)TestCase \{", text, re.S):
        name, body = m.group(1), m.group(2)
        lm = re.search(r"let labels = vec!\[(.*?)\]\s*;", body, re.S)
        cm = re.search(r"let coeffs = vec!\[(.*?)\]\s*;", body, re.S)
        if not lm or not cm:
            continue
        labels = re.findall(r'"([^"]*)"', lm.group(1))
        coeffs = [complex(s) for s in re.findall(r'"([^"]*)"', cm.group(1))]
        if len(labels) != len(coeffs):
            # H10_first100 in fixtures.rs has 100 labels but a longer coeff list;
            # it is not used here.
            continue
        out[name] = (labels, coeffs)
    return out


def parse_precond_vectors(text: str):
    """The dx, e and printed result rv of pyqrusty/tests/test_it.py's precond
    tests (module-level ``dx=np.array([...])``, ``e=...``, ``rv=np.array([...])``).
    Complex literals are parsed with complex(); nothing is executed."""
    def array(name):
        m = re.search(r"^" + name + r"=np\.array\(\[(.*?)\]", text, re.S | re.M)
        return [complex(t.strip().replace(" ", "")) for t in m.group(1).split(",") if t.strip()]
    m = re.search(r"^e=([-+0-9.eE]+)\s*\+\s*0j", text, re.M)
    tol = float(re.search(r"^tol=([-+0-9.eE]+)", text, re.M).group(1))
    return {"dx": array("dx"), "e": float(m.group(1)), "rv": array("rv"), "tol": tol}


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def csr_digests(labels, coeffs, check_dense: bool):
    p = orc.Params(labels, coeffs)
    indptr, indices, data = orc.build_csr(p, step=100)
    if check_dense:
        import scipy.sparse as sps
        A = sps.csr_matrix((data, indices.astype(np.int64), indptr.astype(np.int64)),
                           shape=(p.N, p.N))
        dense = orc.dense_sum(labels, coeffs, "rowwise")
        if not np.array_equal(A.toarray(), dense):
            raise SystemExit("oracle CSR != dense Kronecker sum: refusing to write goldens")
    G = len(np.unique(p.x))
    return {
        "num_qubits": p.num_qubits, "T": p.T, "G": G, "nnz": int(len(indices)),
        "indptr_sha256": digest(indptr), "indices_sha256": digest(indices),
        "data_sha256": digest(data),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    args = ap.parse_args()
    with open(os.path.join(args.reference, "pyqrusty/tests/H_fixtures.py")) as f:
        py = parse_py_fixtures(f.read())
    with open(os.path.join(args.reference, "qrusty/src/fixtures.rs")) as f:
        rs = parse_rs_fixtures(f.read())

    ham = {}
    for n in PY_FIXTURES:
        labels, coeffs = py[n]
        ham[f"py_{n}"] = {"labels": labels, "coeffs": [[c.real, c.imag] for c in coeffs]}
    for n in RS_FIXTURES:
        labels, coeffs = rs[n]
        ham[f"rs_{n}"] = {"labels": labels, "coeffs": [[c.real, c.imag] for c in coeffs]}
    with open(os.path.join(HERE, "hamiltonians.json"), "w") as f:
        json.dump(ham, f, separators=(",", ":"))

    dig = {}
    for name, h in ham.items():
        labels = h["labels"]
        coeffs = [complex(a, b) for a, b in h["coeffs"]]
        nq = len(labels[0])
        if nq > 16:
            continue
        dig[name] = csr_digests(labels, coeffs, check_dense=(nq <= 12))
        print(name, dig[name]["G"], dig[name]["nnz"])
    with open(os.path.join(HERE, "csr_digests.json"), "w") as f:
        json.dump(dig, f, indent=1)

    # the precond inputs and printed result of pyqrusty/tests/test_it.py:276-300
    with open(os.path.join(args.reference, "pyqrusty/tests/test_it.py")) as f:
        pc = parse_precond_vectors(f.read())
    with open(os.path.join(HERE, "precond_h2.json"), "w") as f:
        json.dump({"dx": [[c.real, c.imag] for c in pc["dx"]], "e": pc["e"], "tol": pc["tol"],
                   "rv_printed": [[c.real, c.imag] for c in pc["rv"]]}, f, indent=1)

    # the CPU-runnable workload: 12 qubits, 200 terms
    labels, coeffs = workloads.random_op(12, 200, seed=workloads.SEED_CFG1)
    x = workloads.random_states(1 << 12, 1, seed=workloads.SEED_CFG1 + 1)[0]
    d = csr_digests(labels, coeffs, check_dense=True)
    p = orc.Params(labels, coeffs)
    indptr, indices, data = orc.build_csr(p, step=100)
    y = orc.spmat_dot_densevec(indptr, indices, data, x)
    cfg = {"labels": labels, "coeffs": [[c.real, c.imag] for c in coeffs],
           "x_sha256": digest(x), "csr": d}
    with open(os.path.join(HERE, "config1.json"), "w") as f:
        json.dump(cfg, f, separators=(",", ":"))
    np.save(os.path.join(HERE, "config1_y.npy"), y)
    print("config1", d["G"], d["nnz"])


if __name__ == "__main__":
    main()
