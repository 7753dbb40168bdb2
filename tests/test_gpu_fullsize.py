"""The fp-shell 4p4n M=0 system at full size (n = 1,963,461, seeded synthetic interaction,
626 M upper entries), through checks that hold at any size:

* row windows of the GPU-built H (start, middle, end of the basis and random places) are
  bit-exact against the oracle's rows built from the same m-scheme tables;
* a product with unit vectors returns whole columns of H, whose part on and after the
  diagonal is the oracle's upper row bit for bit in FP64, and its single rounding in FP32;
* the stored product is symmetric (y'Hx = x'Hy) and linear, the FP32 product is within
  FP32 rounding of the FP64 one, and the product form (no H in memory) gives the same
  FP64 products;
* the diagonal read back equals the unit-vector products' diagonal entries;
* the 10-pair mixed-precision solve is within 1 keV of the all-FP64 one, with residuals
  at most 1e-6 as the solver reports them.
"""
import numpy as np
import pytest

import oracle
from cosmo_amd import synth

pytestmark = pytest.mark.gpu

N_FULL = 1963461
WINDOW = 2000


@pytest.fixture(scope="module")
def full(gpu, tmp_path_factory):
    d = tmp_path_factory.mktemp("fp4p4n")
    xml = synth.write_system(str(d), "fp", "fp4p4n", 4, 4, 0, "+", 2)
    s = gpu.System.from_xml(xml)
    assert s.dim == N_FULL
    print("fp 4p4n: building H", flush=True)
    op = gpu.Operator.build(s)
    op.set_layout(2)  # the tiled layout from the first product (the library builds it after 256)
    o = oracle.System(xml)
    b = o.basis()
    yield s, op, o, b
    op.close()
    s.close()


def windows(n):
    rng = np.random.default_rng(5)
    starts = [0, n // 2, n - WINDOW] + sorted(int(x) for x in rng.integers(0, n - WINDOW, 8))
    return [(r, r + WINDOW) for r in starts]


def test_basis_matches_oracle(full):
    s, _, _, b = full
    assert np.array_equal(s.basis(), b)


def test_row_windows_bit_exact_vs_oracle(full):
    _, op, o, b = full
    for r0, r1 in windows(N_FULL):
        rp, c, v = op.upper_rows(r0, r1)
        orp, oc, ov = o.h_upper(b, r0=r0, r1=r1)
        assert np.array_equal(rp, orp), "row lengths differ in rows %d..%d" % (r0, r1)
        assert np.array_equal(c, oc), "columns differ in rows %d..%d" % (r0, r1)
        assert np.array_equal(v.view(np.uint64), ov.view(np.uint64)), "values differ in rows %d..%d" % (r0, r1)


def test_unit_vector_products_are_oracle_rows(gpu, full):
    _, op, o, b = full
    rng = np.random.default_rng(7)
    picks = [0, N_FULL - 1] + [int(x) for x in rng.integers(0, N_FULL, 14)]
    X = np.zeros((N_FULL, len(picks)))
    for i, j in enumerate(picks):
        X[j, i] = 1.0
    Y64 = op.apply(X, gpu.PREC_FP64)
    Y32 = op.apply(X, gpu.PREC_FP32)
    d = op.diagonal()
    for i, j in enumerate(picks):
        orp, oc, ov = o.h_upper(b, r0=j, r1=j + 1)
        # the diagonal read back is the product's own entry and the oracle's value
        assert d[j] == Y64[j, i]
        assert d[j] == (ov[0] if len(oc) and oc[0] == j else 0.0)
        tail = np.nonzero(Y64[j:, i])[0] + j
        assert np.array_equal(tail, oc), "column %d: pattern differs from the oracle's row" % j
        assert np.array_equal(Y64[oc, i].view(np.uint64), ov.view(np.uint64)), "column %d: FP64 values" % j
        assert np.array_equal(Y32[oc, i], ov.astype(np.float32).astype(np.float64)), "column %d: FP32 values" % j
        assert np.array_equal(np.nonzero(Y32[:, i])[0], np.nonzero(Y64[:, i])[0])


def test_products_symmetric_linear_and_consistent(gpu, full):
    s, op, _, _ = full
    rng = np.random.default_rng(11)
    X = rng.standard_normal((N_FULL, 4))
    Y = op.apply(X, gpu.PREC_FP64)
    scale = np.max(np.abs(Y))
    # symmetry: x_a' H x_b = x_b' H x_a for every pair
    G = X.T @ Y
    assert np.max(np.abs(G - G.T)) <= 1e-12 * np.max(np.abs(G))
    # linearity: H (2 x0 - 3 x1) = 2 H x0 - 3 H x1
    Z = op.apply(2.0 * X[:, :1] - 3.0 * X[:, 1:2], gpu.PREC_FP64)
    assert np.max(np.abs(Z[:, 0] - (2.0 * Y[:, 0] - 3.0 * Y[:, 1]))) <= 1e-12 * 5 * scale
    # FP32 values, vectors and accumulation: relative error of a 640-term sum of FP32 products
    Y32 = op.apply(X, gpu.PREC_FP32)
    assert np.max(np.abs(Y32 - Y)) <= 2e-5 * scale
    # the product form builds no H and must give the same FP64 products
    fop = gpu.Operator.build(s, gpu.OP_FACTORED)
    try:
        F = fop.apply(X, gpu.PREC_FP64)
    finally:
        fop.close()
    assert np.max(np.abs(F - Y)) <= 1e-12 * scale


def test_ten_pair_solve_mixed_vs_fp64(gpu, full):
    """The 10-pair solve (block 16, residual 1e-6) on the product form: the mixed-precision
    eigenvalues within 1 keV of the all-FP64 ones, and every returned pair's residual
    recomputed here with the stored FP64 product, independently of the solver."""
    s, op, _, _ = full
    fop = gpu.Operator.build(s, gpu.OP_FACTORED)
    try:
        kw = dict(mblock=16, lim=60, critr=1e-6)
        e64, _, r64, _, st64 = gpu.davidson(fop, 10, want_vectors=False, precision=gpu.PREC_FP64, **kw)
        e32, _, r32, V, st32 = gpu.davidson(fop, 10, want_vectors=True, precision=gpu.PREC_FP32, **kw)
    finally:
        fop.close()
    assert st64["converged"] and st32["converged"]
    assert np.max(r64) <= 1e-6 and np.max(r32) <= 1e-6
    assert np.all(np.diff(e64) >= 0) and np.all(np.diff(e32) >= 0)
    assert np.max(np.abs(e32 - e64)) <= 1e-3  # MeV: 1 keV
    assert st32["matvecs_fp64"] < st32["matvecs"]  # the mixed solve ran an FP32 phase
    HV = op.apply(V, gpu.PREC_FP64)
    assert np.max(np.abs(V.T @ V - np.eye(10))) <= 1e-8
    res = np.linalg.norm(HV - V * e32[None, :], axis=0)
    assert np.max(res) <= 2e-6, res
