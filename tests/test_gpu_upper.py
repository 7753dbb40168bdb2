"""Upper-triangle rows -> the operator's full symmetric rows (cosmo_op_create_upper,
cosmo_op_read_hh, cosmo_op_read_csb): assembled on the GPU, with the host transposition
kept for columns too long for the GPU's LDS sort."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _random_upper(n, per_row, seed, tiny=0):
    rng = np.random.default_rng(seed)
    rp, cols, vals = [0], [], []
    for r in range(n):
        c = np.unique(rng.integers(r + 1, n, size=per_row)) if r + 1 < n else np.zeros(0, np.int64)
        c = np.concatenate([[r], c]).astype(np.int64)
        v = rng.standard_normal(len(c))
        if tiny and r % 7 == 0 and len(c) > 1:
            v[1] = 1e-15  # an off-diagonal entry below the 1e-14 drop
        if tiny and r % 11 == 0:
            v[0] = 1e-16  # a tiny diagonal entry: always kept
        cols.append(c)
        vals.append(v)
        rp.append(rp[-1] + len(c))
    return np.array(rp, np.int64), np.concatenate(cols), np.concatenate(vals)


def _dense(n, rp, c, v, drop):
    H = np.zeros((n, n))
    for r in range(n):
        for e in range(rp[r], rp[r + 1]):
            if drop and c[e] != r and abs(v[e]) < 1e-14:
                continue
            H[r, c[e]] = v[e]
            H[c[e], r] = v[e]
    return H


@pytest.mark.parametrize("flags_name", ["OP_STORED", "OP_KEEP_SMALL"])
def test_gpu_assembly_matches_dense(gpu, flags_name):
    """Random sparse upper rows with entries below 1e-14: the product equals the dense
    product (dropped or kept as the flag says; a tiny diagonal stays); the rows are full,
    ascending, and the diagonal is the dense diagonal."""
    n = 1500
    rp, c, v = _random_upper(n, 12, 3, tiny=1)
    flags = getattr(gpu, flags_name)
    op = gpu.Operator.from_upper(n, rp, c, v, flags)
    H = _dense(n, rp, c, v, drop=flags_name == "OP_STORED")
    X = np.random.default_rng(1).standard_normal((n, 3))
    y = op.apply(X, gpu.PREC_FP64)
    assert np.max(np.abs(y - H @ X)) < 1e-12 * np.abs(H @ X).max()
    assert np.array_equal(op.diagonal(), np.diag(H))
    # the stored upper rows are the input rows less the dropped entries, in input order
    keep = np.ones(len(c), bool)
    if flags_name == "OP_STORED":
        rows = np.repeat(np.arange(n), np.diff(rp))
        keep = (c == rows) | (np.abs(v) >= 1e-14)
    urp, uc, uv = op.upper_rows()
    assert np.array_equal(uc, c[keep]) and np.array_equal(uv, v[keep])
    assert op.nnz()[0] == urp[-1] == keep.sum()


def test_long_column_uses_host_path(gpu):
    """A column with more lower entries than the GPU sort holds (every row couples to the
    last state) goes through the host transposition; same product as the dense matrix."""
    n = 6000
    rng = np.random.default_rng(5)
    rp, cols, vals = [0], [], []
    for r in range(n):
        c = [r] if r == n - 1 else [r, n - 1]
        cols.append(np.array(c, np.int64))
        vals.append(rng.standard_normal(len(c)))
        rp.append(rp[-1] + len(c))
    rp, c, v = np.array(rp, np.int64), np.concatenate(cols), np.concatenate(vals)
    op = gpu.Operator.from_upper(n, rp, c, v)
    X = rng.standard_normal((n, 2))
    H = _dense(n, rp, c, v, drop=True)
    y = op.apply(X, gpu.PREC_FP64)
    assert np.max(np.abs(y - H @ X)) < 1e-12 * np.abs(H @ X).max()


def test_upper_rows_must_ascend(gpu):
    """A column below its row, out of order or repeated is an error, not a silent
    reordering."""
    bad = [
        ([0, 1, 3, 4], [0, 0, 1, 2]),  # row 1 holds column 0
        ([0, 3, 4, 5], [0, 2, 1, 1, 2]),  # row 0: columns 2 then 1
        ([0, 3, 4, 5], [0, 1, 1, 1, 2]),  # row 0: column 1 twice
    ]
    for rp, c in bad:
        rp, c = np.array(rp, np.int64), np.array(c, np.int64)
        with pytest.raises(gpu.CosmoError):
            gpu.Operator.from_upper(3, rp, c, np.ones(len(c)))


def test_large_upload_through_pinned_staging(gpu):
    """Upper rows larger than one 32 MB staging buffer (5 M entries: 40 MB of columns and
    40 MB of values) go through the pinned double buffers; the product equals the same
    symmetric matrix's product in scipy to 1e-12."""
    import scipy.sparse as sp
    n, per = 200000, 50
    rng = np.random.default_rng(8)
    c = np.sort(rng.integers(0, n, size=(n, per)), axis=1)
    rows = np.repeat(np.arange(n), per)
    c = c.ravel()
    keep = c > rows  # strictly upper, then the diagonal added per row
    r_u, c_u = rows[keep], c[keep]
    A = sp.coo_matrix((rng.standard_normal(len(r_u)), (r_u, c_u)), shape=(n, n)).tocsr()
    A.sum_duplicates()
    A = A + sp.diags(rng.standard_normal(n))
    A = sp.csr_matrix(A)
    A.sort_indices()
    assert A.nnz > 4_200_000
    op = gpu.Operator.from_upper(n, A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data,
                                 gpu.OP_KEEP_SMALL)
    H = A + sp.triu(A, k=1).T
    X = rng.standard_normal((n, 3))
    ref = H @ X
    y = op.apply(X, gpu.PREC_FP64)
    assert np.max(np.abs(y - ref)) < 1e-12 * np.abs(ref).max()


def test_read_hh_rejects_bad_files(gpu, golden, tmp_path):
    """A .HH file cut short, with a row count beyond its size, or with bytes after the last
    row is an error."""
    g = golden("sd2p2n_s11")
    op = gpu.Operator.build(gpu.System.from_xml(g.xml), gpu.OP_KEEP_SMALL)
    p = str(tmp_path / "h.HH")
    op.write_hh(p)
    good = open(p, "rb").read()
    bad = {"cut": good[:-5], "trail": good + b"\0" * 16,
           "rows": np.array([10 ** 9], "<u8").tobytes() + good[8:]}
    for k, data in bad.items():
        q = str(tmp_path / (k + ".HH"))
        open(q, "wb").write(data)
        with pytest.raises(gpu.CosmoError):
            gpu.Operator.read_hh(q)
    assert gpu.Operator.read_hh(p).n == op.n
