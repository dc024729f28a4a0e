"""GPU parity of the CSR SpMM kernel (mmsb_spmm_f32) against the oracle's FP64 row loop."""
import numpy as np
import pytest

from oracle import csr_spmm

pytestmark = pytest.mark.gpu


def random_csr(rng, m, n, density, empty_rows=()):
    dense = rng.random((m, n)) < density
    for i in empty_rows:
        dense[i] = False
    rowptr = np.zeros(m + 1, np.int32)
    rowptr[1:] = np.cumsum(dense.sum(1))
    colidx = np.nonzero(dense)[1].astype(np.int32)
    return rowptr, colidx


@pytest.mark.parametrize("r", [1, 3, 4, 8, 20, 64, 110, 256, 300, 510, 1010])
@pytest.mark.parametrize("weighted", [False, True])
def test_spmm_matches_oracle(mmsb, dev, r, weighted):
    import torch
    rng = np.random.default_rng(r * 7 + weighted)
    m, n = 257, 190
    rowptr, colidx = random_csr(rng, m, n, 0.08, empty_rows=(0, 5, 256))
    # a long row exercising several 64-index chunks
    vals = rng.standard_normal(colidx.size).astype(np.float32) if weighted else None
    ld = (r + 3) // 4 * 4 + (4 if r % 2 else 0)
    B = np.zeros((n, ld), np.float32)
    B[:, :r] = rng.standard_normal((n, r))
    Y = torch.full((m, ld), 7.0, dtype=torch.float32, device=dev)
    t = lambda a: torch.from_numpy(a).to(dev)
    bufs = [t(rowptr), t(colidx), t(vals) if weighted else None, t(B)]
    mmsb.spmm_f32(*bufs, Y, r=r, n=n, alpha=0.5)
    torch.cuda.synchronize()
    ref = csr_spmm(rowptr, colidx, vals, B[:, :r].astype(np.float64), m, alpha=0.5)
    got = Y.cpu().numpy()
    np.testing.assert_allclose(got[:, :r], ref, rtol=1e-5, atol=1e-5)
    assert np.all(got[:, r:] == 7.0), "columns beyond r must not be written"


def test_spmm_long_rows(mmsb, dev):
    import torch
    rng = np.random.default_rng(11)
    m, n, r = 9, 5000, 100
    rowptr, colidx = random_csr(rng, m, n, 0.6)  # ~3000 nnz per row
    B = rng.standard_normal((n, r)).astype(np.float32)
    Y = torch.empty((m, r), dtype=torch.float32, device=dev)
    t = lambda a: torch.from_numpy(a).to(dev)
    bufs = [t(rowptr), t(colidx), None, t(B)]
    mmsb.spmm_f32(*bufs, Y, r=r, n=n)
    torch.cuda.synchronize()
    np.testing.assert_allclose(Y.cpu().numpy(), csr_spmm(rowptr, colidx, None, B.astype(np.float64), m), rtol=1e-4, atol=1e-3)


def test_spmm_row_window(mmsb, dev):
    """A sub-range of rows via an offset rowptr (used by checkpointed STGD)."""
    import torch
    rng = np.random.default_rng(5)
    m, n, r = 100, 80, 16
    rowptr, colidx = random_csr(rng, m, n, 0.1)
    B = rng.standard_normal((n, r)).astype(np.float32)
    # every device buffer stays referenced until after the synchronize below
    d_rowptr = torch.from_numpy(rowptr).to(dev)
    d_colidx = torch.from_numpy(colidx).to(dev)
    d_B = torch.from_numpy(B).to(dev)
    Y = torch.empty((10, r), dtype=torch.float32, device=dev)
    assert int(rowptr[50]) <= colidx.size and colidx.max() < n
    st = mmsb.lib.mmsb_spmm_f32(None, 10, n, int(colidx.size), d_rowptr.data_ptr() + 4 * 40,
                                d_colidx.data_ptr(), None, 1.0, d_B.data_ptr(), r, r, Y.data_ptr(), r)
    mmsb.check(st)
    torch.cuda.synchronize()
    full = csr_spmm(rowptr, colidx, None, B.astype(np.float64), m)
    np.testing.assert_allclose(Y.cpu().numpy(), full[40:50], rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("r", [20, 510])
def test_row_order_and_ordered_spmm(mmsb, dev, r):
    """mmsb_csr_row_order sorts rows by descending length with ties in row order (compared with a
    stable numpy argsort), and the SpMM that visits rows in that order writes the same Y bit for bit."""
    import torch
    rng = np.random.default_rng(40 + r)
    m, n = 3000, 800
    lens = np.minimum(rng.pareto(1.5, m) * 3, n - 1).astype(np.int64)
    lens[[0, 17, 2999]] = 0                                   # empty rows, first and last included
    rowptr = np.zeros(m + 1, np.int32)
    rowptr[1:] = np.cumsum(lens)
    colidx = np.concatenate([np.sort(rng.choice(n, size=int(l), replace=False)) for l in lens]).astype(np.int32)
    B = rng.standard_normal((n, (r + 3) // 4 * 4)).astype(np.float32)
    t = lambda a: torch.from_numpy(a).to(dev)
    d_rp, d_ci, d_B = t(rowptr), t(colidx), t(B)
    order = mmsb.csr_row_order(d_rp)
    torch.cuda.synchronize()
    assert np.array_equal(order.cpu().numpy(), np.argsort(-lens, kind="stable"))
    Y0 = torch.empty((m, B.shape[1]), dtype=torch.float32, device=dev)
    Y1 = torch.empty_like(Y0)
    mmsb.spmm_f32(d_rp, d_ci, None, d_B, Y0, r=r, n=n)
    mmsb.spmm_f32(d_rp, d_ci, None, d_B, Y1, r=r, n=n, order=order)
    torch.cuda.synchronize()
    assert torch.equal(Y0[:, :r], Y1[:, :r])
    np.testing.assert_allclose(Y1[:, :r].cpu().numpy(), csr_spmm(rowptr, colidx, None, B[:, :r].astype(np.float64), m),
                               rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("chunk_nnz", [1, 50, 700, 1 << 22])
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("pinned", [False, True])
def test_spmm_with_host_csr_matches_device(mmsb, dev, chunk_nnz, weighted, pinned):
    """The CSR block in host memory, streamed in row chunks (mmsb_spmm_f32_host_csr): Y identical to
    the device-resident product, for chunks of one nonzero (every row its own chunk, rows longer than
    the chunk included), a few hundred, and the whole block; empty rows at both ends; from pageable
    arrays (staged) and from pinned ones (copied directly)."""
    import torch
    rng = np.random.default_rng(chunk_nnz + weighted)
    m, n, r = 1500, 900, 130
    rowptr, colidx = random_csr(rng, m, n, 0.05, empty_rows=(0, 1, 700, 1499))
    vals = rng.uniform(0.5, 2.0, colidx.size).astype(np.float32) if weighted else None
    if pinned:
        def pin(a):
            t = torch.empty(a.size, dtype=torch.int32 if a.dtype == np.int32 else torch.float32, pin_memory=True)
            t.numpy()[:] = a
            return t.numpy()
        colidx = pin(colidx)
        vals = None if vals is None else pin(vals)
    B = torch.from_numpy(rng.standard_normal((n, 132)).astype(np.float32)).to(dev)
    want = torch.full((m, 132), 3.0, dtype=torch.float32, device=dev)
    got = torch.full((m, 132), 3.0, dtype=torch.float32, device=dev)
    t = lambda a: torch.from_numpy(a).to(dev)
    mmsb.spmm_f32(t(rowptr), t(colidx), None if vals is None else t(vals), B, want, r=r, n=n, alpha=0.25)
    mmsb.spmm_f32_host_csr(rowptr, colidx, vals, B, got, r=r, n=n, alpha=0.25, chunk_nnz=chunk_nnz)
    torch.cuda.synchronize()
    assert torch.equal(got, want)
    ref = csr_spmm(rowptr, colidx, vals, B[:, :r].double().cpu().numpy(), m, alpha=0.25)
    np.testing.assert_allclose(got[:, :r].cpu().numpy(), ref, rtol=1e-4, atol=1e-4)
    assert (got[:, r:] == 3.0).all()            # columns beyond r untouched


def test_spmm_host_csr_rejects_bad_blocks(mmsb, dev):
    """Host-side checks before any upload: a column index out of range and a decreasing rowptr."""
    import torch
    B = torch.zeros((10, 4), dtype=torch.float32, device=dev)
    Y = torch.zeros((3, 4), dtype=torch.float32, device=dev)
    rowptr = np.array([0, 1, 2, 3], np.int32)
    with pytest.raises(mmsb.MMSBError, match="column index"):
        mmsb.spmm_f32_host_csr(rowptr, np.array([0, 10, 2], np.int32), None, B, Y, r=4, n=10)
    with pytest.raises(mmsb.MMSBError, match="decreases"):
        mmsb.spmm_f32_host_csr(np.array([0, 2, 1, 3], np.int32), np.array([0, 1, 2], np.int32), None, B, Y, r=4, n=10)
