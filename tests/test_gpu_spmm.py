"""H.X on the GPU: ordered FP64 is bit-identical to the reference CSB product order;
fast FP64 / FP32 / virtual within stated tolerances."""
import hashlib

import numpy as np
import pytest

import oracle
from conftest import FULL_CASES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", FULL_CASES + ["sd4p4n_s0"])
def test_ordered_fp64_bit_exact_vs_csb_product(gpu, golden, name):
    g = golden(name)
    s = gpu.System.from_xml(g.xml)
    o = oracle.System(g.xml)
    op = gpu.Operator.build(s)
    rp, c, v = o.h_upper(drop_small=True)
    n = op.n
    X = np.sin(0.37 * np.arange(n)[:, None] + 1.3 * np.arange(3)[None, :])
    Yo = oracle.csb_times(n, rp, c, v, X)
    Yg = op.apply(X, gpu.PREC_FP64_ORDERED)
    assert np.array_equal(Yg.view(np.uint64), Yo.view(np.uint64))


@pytest.mark.parametrize("name", FULL_CASES)
def test_reference_product_on_reference_values(gpu, golden, name):
    """The reference's own TimesMatrix output (oracle/_ref harness) on its own H: the GPU
    operator uploaded from the reference .HH rows reproduces it bit for bit."""
    g = golden(name)
    a = g.arrays
    n = g.meta["n"]
    op = gpu.Operator.from_upper(n, a["hh_rowptr"], a["hh_cols"], a["hh_vals"])  # drops |h|<1e-14 like CSB
    Y = op.apply(a["spmm_x"], gpu.PREC_FP64_ORDERED)
    assert np.array_equal(Y.view(np.uint64), a["spmm_y"].view(np.uint64))


@pytest.mark.parametrize("m", [1, 3, 8, 10, 16])
def test_fast_fp64_and_fp32(gpu, golden, m):
    g = golden("sd4p4n_s0")
    op = gpu.Operator.build(gpu.System.from_xml(g.xml))
    o = oracle.System(g.xml)
    rp, c, v = o.h_upper(drop_small=True)
    n = op.n
    X = np.random.default_rng(m).standard_normal((n, m))
    ref = oracle.csb_times(n, rp, c, v, X)
    scale = np.abs(ref).max()
    y64 = op.apply(X, gpu.PREC_FP64)
    assert np.max(np.abs(y64 - ref)) < 1e-12 * scale
    y32 = op.apply(X, gpu.PREC_FP32)
    assert np.max(np.abs(y32 - ref)) < 2e-5 * scale


@pytest.mark.parametrize("name", ["sd4p4n_s0", "fp2p2n_s5"])
def test_virtual_equals_stored(gpu, golden, name):
    g = golden(name)
    s = gpu.System.from_xml(g.xml)
    a = gpu.Operator.build(s)
    b = gpu.Operator.build(s, gpu.OP_VIRTUAL)
    n = a.n
    X = np.random.default_rng(7).standard_normal((n, 5))
    ya = a.apply(X, gpu.PREC_FP64)
    yb = b.apply(X, gpu.PREC_FP64)
    assert np.max(np.abs(ya - yb)) < 1e-12 * np.abs(ya).max()
    yb32 = b.apply(X, gpu.PREC_FP32)
    assert np.max(np.abs(yb32 - ya)) < 2e-5 * np.abs(ya).max()


def test_apply_device_torch(gpu, golden):
    import torch
    g = golden("sd4p4n_s0")
    op = gpu.Operator.build(gpu.System.from_xml(g.xml))
    n = op.n
    X = torch.randn(n, 16, dtype=torch.float64, device="cuda")
    Y = torch.empty_like(X)
    op.apply_device(X, Y, gpu.PREC_FP64, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    ref = op.apply(X.cpu().numpy(), gpu.PREC_FP64)
    assert np.max(np.abs(Y.cpu().numpy() - ref)) < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("name", ["sd2p2n_s11", "sd3p1n_m1_s7", "fp2p2n_s5", "fp3p2n_m1_s3", "sd4p4n_s0"])
def test_tiled_layout_vs_oracle_and_rows(gpu, golden, name):
    """The tiled stored product (row blocks of whole neutron configurations, X tiles in
    LDS, used by default from n = 262144 after 256 products) forced on the golden systems: FP64 within 1e-12
    and FP32 within 2e-5 of the oracle's CSB product for every row pitch it serves, the
    same bits on a repeated call, and the row kernel's result to rounding."""
    g = golden(name)
    s = gpu.System.from_xml(g.xml)
    o = oracle.System(g.xml)
    rp, c, v = o.h_upper(drop_small=True)
    tiled = gpu.Operator.build(s)
    tiled.set_layout(2)
    rows = gpu.Operator.build(s)
    rows.set_layout(1)
    n = tiled.n
    for m in (1, 2, 3, 4, 7, 8, 12, 16):
        X = np.random.default_rng(m).standard_normal((n, m))
        ref = oracle.csb_times(n, rp, c, v, X)
        scale = np.abs(ref).max()
        y64 = tiled.apply(X, gpu.PREC_FP64)
        y32 = tiled.apply(X, gpu.PREC_FP32)
        assert np.abs(y64 - ref).max() <= 1e-12 * scale, m
        assert np.abs(y32 - ref).max() <= 2e-5 * scale, m
        assert np.array_equal(tiled.apply(X, gpu.PREC_FP64), y64)
        assert np.abs(rows.apply(X, gpu.PREC_FP64) - y64).max() <= 1e-12 * scale


def test_tiled_layout_davidson(gpu, golden):
    """Davidson on the tiled product (FP64 and mixed) equals the row kernel's solve."""
    g = golden("sd4p4n_s0")
    s = gpu.System.from_xml(g.xml)
    a = gpu.Operator.build(s)
    a.set_layout(2)
    b = gpu.Operator.build(s)
    b.set_layout(1)
    for prec in (gpu.PREC_FP64, gpu.PREC_FP32):
        ea, _, ra, _, sa = gpu.davidson(a, 10, want_vectors=False, critr=1e-6, lim=40, precision=prec)
        eb, _, rb, _, sb = gpu.davidson(b, 10, want_vectors=False, critr=1e-6, lim=40, precision=prec)
        assert sa["converged"] and sb["converged"]
        assert np.abs(ea - eb).max() < 1e-9 and ra.max() <= 1e-6


@pytest.mark.parametrize("name", ["psd2p2n_hw2", "fp3p2n_m1_rej", "fp2p2n_s5_jj", "rows:fp2p2n_s5"])
def test_tiled_layout_without_the_proton_spectator_split(gpu, golden, name):
    """The tiled product forced where its proton-spectator split does not apply (a truncated
    basis; an operator uploaded from upper rows, which carries no basis structure: blocks of
    576 rows) or where added JJ terms widen each pair's set of values: FP64 within 1e-12 and
    FP32 within 2e-5 of the CSB product of the oracle's rows (of the rows the reference's
    XHH+JJ wrote, for the JJ system) at every row pitch it serves."""
    from conftest import TRUNC_CASES, trunc_case
    from_rows = name.startswith("rows:")
    name = name.split(":")[-1]
    xml = trunc_case(name)[0] if name in TRUNC_CASES else golden(name).xml
    s = gpu.System.from_xml(xml)
    if name.endswith("_jj"):
        a = golden(name).arrays
        rp, c, v = a["hh_rowptr"], a["hh_cols"], a["hh_vals"]
    else:
        rp, c, v = oracle.System(xml).h_upper(basis=s.basis(), drop_small=True)
    n = len(rp) - 1
    if from_rows:
        op = gpu.Operator.from_upper(n, np.asarray(rp, np.int64), np.asarray(c, np.int64), np.asarray(v))
    else:
        op = gpu.Operator.build(s)
    op.set_layout(2)
    assert op.n == n
    for m in (1, 3, 4, 8, 16):
        X = np.random.default_rng(10 + m).standard_normal((n, m))
        ref = oracle.csb_times(n, rp, c, v, X)
        scale = np.abs(ref).max()
        assert np.abs(op.apply(X, gpu.PREC_FP64) - ref).max() <= 1e-12 * scale, m
        assert np.abs(op.apply(X, gpu.PREC_FP32) - ref).max() <= 2e-5 * scale, m
