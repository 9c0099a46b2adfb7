"""GPU parity of the batched build + matvec (qr_batch_run).

Every operator of a batch gets its CSR bit-identical to the oracle's rowwise
build (all rows for small operators, sampled rows otherwise) and y_k = A_k x_k
within 1e-12 relative (normwise, and componentwise against
sum_t |c'_t||x[r ^ x_t]|).  The batch mixes every register class of the
persistent kernel (1, 2, 4 and 6 group chunks per thread) with an operator of
more than 6144 distinct x masks, which runs through the single-operator
kernels in the same graph.  Graph replays with new states, every tuning
candidate of the fused form and the y-only mode must give bit-identical
results; the split form (build pass, then a matvec pass over the written CSR)
the same CSR and y within 1e-13.
"""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# (qubits, terms, seed): G = distinct x masks decides the kernel class
SPECS = [(8, 10, 1), (13, 2600, 2), (1, 2, 3), (12, 1500, 4), (6, 40, 5),
         (13, 7000, 6), (11, 300, 7), (14, 4200, 8), (2, 3, 9), (10, 1, 10), (14, 9000, 11)]
FULL_LIMIT = 1 << 22   # entries; above this the CSR is checked on sampled rows


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import qrusty_amd  # noqa: F401
    yield
    torch.cuda.synchronize()


def _abs_bound(p, x, rows):
    cabs = np.abs(p.cp.view(np.complex128))
    out = np.zeros(len(rows))
    for t in range(p.T):
        out += cabs[t] * np.abs(x[(rows ^ p.x[t]).astype(np.int64)])
    return out


@pytest.fixture(scope="module")
def batch_case():
    import qrusty_amd as qa
    from qrusty_amd import workloads
    ops, params = [], []
    for nq, nt, seed in SPECS:
        labels, coeffs = workloads.random_op(nq, nt, seed=0x5150B000 + seed)
        labels = [("-i" if k % 7 == 3 else "") + l for k, l in enumerate(labels)]
        ops.append(qa.SparsePauliOp([qa.Pauli(l) for l in labels], coeffs))
        params.append(orc.Params(labels, coeffs))
    Gs = [len(np.unique(p.x)) for p in params]
    print(f"[gpu] batch specs {SPECS} G={Gs}", flush=True)
    # every class of the persistent kernel, and the single-operator route
    assert any(G <= 1024 for G in Gs) and any(1024 < G <= 2048 for G in Gs)
    assert any(2048 < G <= 4096 for G in Gs) and any(4096 < G <= 6144 for G in Gs)
    assert any(G > 6144 for G in Gs)
    return ops, params


def _states(params, seed):
    from qrusty_amd import workloads
    return [workloads.random_states(p.N, 1, seed + k)[0] for k, p in enumerate(params)]


def _check_op(b, k, p, x):
    """CSR bit-exact vs the oracle and y vs the oracle rows."""
    y = b.y[k].cpu().numpy()
    N = p.N
    G = len(np.unique(p.x))
    rows = np.arange(N, dtype=np.uint64) if N * G <= FULL_LIMIT else np.unique(np.concatenate(
        [[0, N - 1], np.random.default_rng(k).choice(N, size=96, replace=False)])).astype(np.uint64)
    csr = b.csr(k)
    if csr is not None:
        indptr = csr[0].cpu().numpy().view(np.uint64)
        assert np.array_equal(indptr, np.arange(N + 1, dtype=np.uint64) * np.uint64(G))
        if len(rows) == N:
            ip, ix, dv = orc.build_csr(p, step=100)
            assert np.array_equal(indptr, ip)
            assert np.array_equal(csr[1].cpu().numpy().view(np.uint64), ix)
            assert np.array_equal(csr[2].cpu().numpy().view(np.float64), dv.view(np.float64))
        else:
            sel = torch.from_numpy((rows[:, None] * G + np.arange(G, dtype=np.uint64)).astype(np.int64)
                                   .reshape(-1)).to(csr[1].device)
            gi = csr[1][sel].cpu().numpy().view(np.uint64).reshape(len(rows), G)
            gd = csr[2][sel].cpu().numpy().reshape(len(rows), G)
            for i, r in enumerate(rows):
                cols, vals = orc.make_row(p, int(r))
                assert np.array_equal(gi[i], cols), f"op {k} row {r} indices"
                assert np.array_equal(gd[i].view(np.float64), vals.view(np.float64)), \
                    f"op {k} row {r} data"
    ref = orc.apply_rows(p, x, rows)[0]
    got = y[rows.astype(np.int64)]
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref) + 1e-300
    assert np.all(np.abs(got - ref) <= 1e-12 * _abs_bound(p, x, rows) + 1e-300), f"op {k}"
    return y


def test_batch_vs_oracle_and_replays(batch_case):
    import qrusty_amd as qa
    ops, params = batch_case
    b = qa.Batch(ops, keep_csr=True)
    ys = []
    for rep in range(3):  # first run records the graph, the others replay it
        xs = _states(params, seed=100 * rep)
        b.set_states(xs)
        b.run()
        torch.cuda.synchronize()
        print(f"[gpu] batch replay {rep}", flush=True)
        ys.append([_check_op(b, k, p, xs[k]) for k, p in enumerate(params)])
        # eager work between replays
        torch.cuda.synchronize()
    # same states again: bit-identical to the first run
    b.set_states(_states(params, seed=0))
    b.run()
    torch.cuda.synchronize()
    for k in range(len(ops)):
        assert np.array_equal(b.y[k].cpu().numpy(), ys[0][k])


@pytest.mark.parametrize("split", [1, 0])
def test_batch_heavy_groups(split):
    """Operators with a heavy diagonal group (60-150 Z strings) beside random
    terms: one of more than 2048 groups (sub-operators in the split form), one
    of 1025-2048 and one of at most 1024 groups.  The split form folds heavy
    groups inline in its sub-operator kernels; CSR bit-exact, y at 1e-12."""
    import qrusty_amd as qa
    rng = np.random.default_rng(4242)
    ops, params = [], []
    for n, T, nd in [(13, 3200, 150), (11, 1500, 60), (12, 500, 100)]:
        labels = ["".join(rng.choice(list("IXYZ"), size=n)) for _ in range(T)]
        labels += ["".join(rng.choice(["I", "Z"], size=n)) for _ in range(nd)]
        coeffs = rng.standard_normal(len(labels)) + 1j * rng.standard_normal(len(labels))
        ops.append(qa.SparsePauliOp([qa.Pauli(l) for l in labels], coeffs))
        params.append(orc.Params(labels, coeffs))
    Gs = [len(np.unique(p.x)) for p in params]
    print(f"[gpu] batch heavy groups split={split} G={Gs}", flush=True)
    assert Gs[0] > 2048 and 1024 < Gs[1] <= 2048 and Gs[2] <= 1024
    b = qa.Batch(ops, keep_csr=True)
    xs = _states(params, seed=31)
    b.set_states(xs)
    b.run({"tiles_per_item": 2, "blocks_per_cu": 2, "order": 0, "split": split})
    torch.cuda.synchronize()
    for k, p in enumerate(params):
        _check_op(b, k, p, xs[k])


def test_batch_tunings_bit_identical(batch_case):
    import qrusty_amd as qa
    ops, params = batch_case
    b = qa.Batch(ops, keep_csr=True)
    b.set_states(_states(params, seed=7))
    b.run({"tiles_per_item": 1, "blocks_per_cu": 1, "order": 0})
    torch.cuda.synchronize()
    y0 = [t.cpu().numpy() for t in b.y]
    c0 = [tuple(a.cpu().numpy() for a in b.csr(k)) for k in range(len(ops))]
    for tpi, bpc, order, split in [(2, 2, 1, 0), (8, 1, 1, 0), (4, 2, 0, 0), (1, 1, 0, 1),
                                   (4, 2, 1, 1)]:
        print(f"[gpu] batch tuning tpi={tpi} bpc={bpc} order={order} split={split}", flush=True)
        for t in b.y:
            t.zero_()
        b.run({"tiles_per_item": tpi, "blocks_per_cu": bpc, "order": order, "split": split})
        torch.cuda.synchronize()
        for k in range(len(ops)):
            for a, ref in zip(b.csr(k), c0[k]):
                assert np.array_equal(a.cpu().numpy(), ref)
            if split == 0:
                assert np.array_equal(b.y[k].cpu().numpy(), y0[k])
            else:  # the matvec pass sums each row in another order
                yk = b.y[k].cpu().numpy()
                assert np.linalg.norm(yk - y0[k]) <= 1e-13 * np.linalg.norm(y0[k]) + 1e-300
    best = b.tune()
    assert best["tiles_per_item"] in (1, 2, 4, 8) and best["blocks_per_cu"] in (1, 2)
    assert best["split"] in (0, 1)
    b.run()
    torch.cuda.synchronize()
    for k in range(len(ops)):
        yk = b.y[k].cpu().numpy()
        assert np.linalg.norm(yk - y0[k]) <= 1e-13 * np.linalg.norm(y0[k]) + 1e-300


def test_batch_y_only(batch_case):
    import qrusty_amd as qa
    ops, params = batch_case
    full = qa.Batch(ops, keep_csr=True)
    yonly = qa.Batch(ops, keep_csr=False)
    xs = _states(params, seed=11)
    full.set_states(xs)
    yonly.set_states(xs)
    full.run()
    yonly.run()
    torch.cuda.synchronize()
    for k, p in enumerate(params):
        if len(np.unique(p.x)) <= 6144:
            assert yonly.csr(k) is None
        assert np.array_equal(yonly.y[k].cpu().numpy(), full.y[k].cpu().numpy())


def test_batch_errors():
    import qrusty_amd as qa
    with pytest.raises(qa.QrError):
        qa.Batch([])
    op = qa.SparsePauliOp([qa.Pauli("XZ")], [1.0])
    b = qa.Batch([op])
    with pytest.raises(qa.QrError):
        b.run({"tiles_per_item": 0, "blocks_per_cu": 1, "order": 0})
    with pytest.raises(qa.QrError):
        b.run({"tiles_per_item": 1, "blocks_per_cu": 9, "order": 0})
    # the split form needs CSR buffers
    yonly = qa.Batch([op], keep_csr=False)
    with pytest.raises(qa.QrError):
        yonly.run({"tiles_per_item": 1, "blocks_per_cu": 1, "order": 0, "split": 1})
