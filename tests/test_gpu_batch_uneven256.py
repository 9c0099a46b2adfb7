"""GPU parity of the full uneven batch: 256 random operators of 8..18 qubits
and 10..5000 terms (log-uniform), every CSR kept (about 213 GB of device
memory), in both forms of qr_batch_run: the matvec fused into the build and
the split form (build pass, then a matvec pass over the written CSR).

The reference has no batch; each operator must equal the one-operator path it
replaces, qrusty's RowwiseUnsafeChunked build and spmat_dot_densevec
(qrusty/src/accel.rs:267-370).  For every operator:
  * indptr equals the closed form r * G on every row;
  * 8 sampled rows (the first, the last and 6 random ones) hold exactly the
    oracle's make_row (accel.rs:171-210) indices and value bits;
  * y at those rows is within 1e-12 of the oracle's rows of A x, normwise and
    componentwise against sum_t |c'_t| |x[r ^ x_t]|.
The two forms give the same CSR bits and y within 1e-13.
"""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ROWS_PER_OP = 8


@pytest.fixture(scope="module")
def uneven256():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import qrusty_amd as qa
    from qrusty_amd import workloads
    specs = workloads.uneven_batch(256, workloads.SEED_CFG4)
    ops = [qa.SparsePauliOp([qa.Pauli(l) for l in lab], co) for lab, co in specs]
    params = [orc.Params(lab, co) for lab, co in specs]
    b = qa.Batch(ops, keep_csr=True)
    xs = [workloads.random_states(p.N, 1, seed=workloads.SEED_CFG4 + k)[0]
          for k, p in enumerate(params)]
    b.set_states(xs)
    nnz = sum(h.nnz for h in b._h)
    print(f"[gpu] uneven batch: 256 operators, nnz={nnz} ({24 * nnz / 1e9:.1f} GB CSR)", flush=True)
    yield b, params, xs
    del b
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _abs_bound(p, x, rows):
    cabs = np.abs(p.cp.view(np.complex128))
    out = np.zeros(len(rows))
    for t in range(p.T):
        out += cabs[t] * np.abs(x[(rows ^ p.x[t]).astype(np.int64)])
    return out


def _check_all(b, params, xs, label):
    ys = []
    for k, p in enumerate(params):
        N, G = p.N, len(np.unique(p.x))
        ip, ix, dv = b.csr(k)
        ref_ip = torch.arange(N + 1, dtype=torch.int64, device=ip.device) * G
        assert torch.equal(ip, ref_ip), f"{label}: op {k} indptr"
        rng = np.random.default_rng(1000 + k)
        rows = np.unique(np.concatenate([[0, N - 1], rng.choice(N, size=ROWS_PER_OP - 2,
                                                               replace=False)])).astype(np.uint64)
        sel = torch.from_numpy((rows[:, None] * G + np.arange(G, dtype=np.uint64))
                               .astype(np.int64).reshape(-1)).to(ix.device)
        gi = ix[sel].cpu().numpy().view(np.uint64).reshape(len(rows), G)
        gd = dv[sel].cpu().numpy().reshape(len(rows), G)
        for i, r in enumerate(rows):
            cols, vals = orc.make_row(p, int(r))
            assert np.array_equal(gi[i], cols), f"{label}: op {k} row {r} indices"
            assert np.array_equal(gd[i].view(np.float64), vals.view(np.float64)), \
                f"{label}: op {k} row {r} values"
        y = b.y[k].cpu().numpy()
        ref = orc.apply_rows(p, xs[k], rows)[0]
        got = y[rows.astype(np.int64)]
        assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref) + 1e-300, f"{label}: op {k} y"
        assert np.all(np.abs(got - ref) <= 1e-12 * _abs_bound(p, xs[k], rows) + 1e-300), \
            f"{label}: op {k} y componentwise"
        ys.append(y)
    return ys


def test_uneven256_fused_and_split(uneven256):
    b, params, xs = uneven256
    print("[gpu] uneven batch: fused form", flush=True)
    b.run({"tiles_per_item": 2, "blocks_per_cu": 2, "order": 1, "split": 0})
    torch.cuda.synchronize()
    y_fused = _check_all(b, params, xs, "fused")
    # wrapping int64 sums of every CSR array (no temporaries: the CSRs fill
    # most of the device memory)
    digests = [tuple(int(a.view(torch.int64).sum().item()) for a in b.csr(k))
               for k in range(len(params))]
    # clear every output, so that the split form's results are its own
    for k in range(len(params)):
        for a in b.csr(k):
            a.zero_()
        b.y[k].zero_()
    print("[gpu] uneven batch: split form", flush=True)
    b.run({"tiles_per_item": 2, "blocks_per_cu": 2, "order": 1, "split": 1})
    torch.cuda.synchronize()
    y_split = _check_all(b, params, xs, "split")
    for k in range(len(params)):
        again = tuple(int(a.view(torch.int64).sum().item()) for a in b.csr(k))
        assert again == digests[k], f"op {k}: CSR differs between the forms"
        assert np.linalg.norm(y_split[k] - y_fused[k]) <= 1e-13 * np.linalg.norm(y_fused[k]) + 1e-300
