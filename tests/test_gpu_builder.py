"""GPU Hamiltonian builder against the oracle (same m-scheme tables, so values are
bit-exact) and against the reference's .HH rows (structure exact, values to CG noise)."""
import numpy as np
import pytest

import oracle
from conftest import FULL_CASES, GOLDEN_CASES

pytestmark = pytest.mark.gpu


def full_rows_from_upper(n, rp, cols, vals):
    """full symmetric rows (cols ascending) from upper rows"""
    rows = [[] for _ in range(n)]
    for r in range(n):
        for e in range(rp[r], rp[r + 1]):
            rows[r].append((cols[e], vals[e]))
            if cols[e] != r:
                rows[cols[e]].append((r, vals[e]))
    return [sorted(x) for x in rows]


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_upper_rows_bit_exact_vs_oracle(gpu, golden, name):
    g = golden(name)
    s = gpu.System.from_xml(g.xml)
    o = oracle.System(g.xml)
    for keep in (False, True):
        op = gpu.Operator.build(s, gpu.OP_KEEP_SMALL if keep else gpu.OP_STORED)
        rp, c, v = op.upper_rows()
        orp, oc, ov = o.h_upper(drop_small=not keep)
        assert np.array_equal(rp, orp), "row lengths differ"
        assert np.array_equal(c, oc), "columns differ"
        assert np.array_equal(v.view(np.uint64), ov.view(np.uint64)), "values not bit-exact"
        nu, nf = op.nnz()
        assert nu == orp[-1]
        if not keep:
            assert nu == g.meta["nnz_upper_csb"]
        else:
            assert nu == g.meta["hh_nnz"]


@pytest.mark.parametrize("name", FULL_CASES)
def test_rows_match_reference_hh(gpu, golden, name):
    g = golden(name)
    op = gpu.Operator.build(gpu.System.from_xml(g.xml), gpu.OP_KEEP_SMALL)
    rp, c, v = op.upper_rows()
    a = g.arrays
    assert np.array_equal(rp, a["hh_rowptr"]) and np.array_equal(c, a["hh_cols"])
    assert np.max(np.abs(v - a["hh_vals"])) < 1e-12


@pytest.mark.parametrize("name", ["sd4p4n_s0", "fp3p2n_m1_s3"])
def test_sampled_rows_match_reference_hh(gpu, golden, name):
    g = golden(name)
    op = gpu.Operator.build(gpu.System.from_xml(g.xml), gpu.OP_KEEP_SMALL)
    a = g.arrays
    for i, r in enumerate(a["sample_rows"]):
        rp, c, v = op.upper_rows(int(r), int(r) + 1)
        s0, s1 = a["sample_rowptr"][i], a["sample_rowptr"][i + 1]
        assert np.array_equal(c, a["sample_cols"][s0:s1])
        assert np.max(np.abs(v - a["sample_vals"][s0:s1])) < 1e-12


def test_hh_file_roundtrip(gpu, golden, tmp_path):
    g = golden("sd2p2n_s11")
    op = gpu.Operator.build(gpu.System.from_xml(g.xml), gpu.OP_KEEP_SMALL)
    p = str(tmp_path / "x.HH")
    op.write_hh(p)
    rp, c, v = oracle.read_hh(p)
    orp, oc, ov = oracle.System(g.xml).h_upper(drop_small=False)
    assert np.array_equal(rp, orp) and np.array_equal(c, oc) and np.array_equal(v, ov)
    op2 = gpu.Operator.read_hh(p)
    assert np.array_equal(op2.diagonal(), op.diagonal())


def test_diagonal_and_symmetry(gpu, golden):
    g = golden("fp2p2n_s5")
    op = gpu.Operator.build(gpu.System.from_xml(g.xml))
    rp, c, v = op.upper_rows()
    d = op.diagonal()
    n = op.n
    assert np.array_equal(d, v[rp[:-1]])  # diagonal first in every row
    assert np.all(c[rp[:-1]] == np.arange(n))
    H = oracle.dense_from_upper(n, rp, c, v)
    X = np.random.default_rng(3).standard_normal((n, 2))
    Y = op.apply(X)
    assert np.allclose(Y, H @ X, rtol=0, atol=1e-11)


def test_virtual_diagonal_equals_stored(gpu, golden):
    g = golden("sd4p4n_s0")
    s = gpu.System.from_xml(g.xml)
    a = gpu.Operator.build(s)
    b = gpu.Operator.build(s, gpu.OP_VIRTUAL)
    assert np.array_equal(a.diagonal(), b.diagonal())


@pytest.mark.parametrize("name", ["sd2p2n_s11_jjtt", "fp2p2n_s5_jj", "sd3p1n_m1_s7_tt"])
def test_jj_tt_rows_match_reference_hh(gpu, golden, name):
    """H with <matrix JJ / TT> add-ons: the rows XHH+JJ wrote (the reference's own .HH), same
    structure and values to the CG noise, on the stored and the product-form operator."""
    g = golden(name)
    s = gpu.System.from_xml(g.xml)
    op = gpu.Operator.build(s, gpu.OP_KEEP_SMALL)
    rp, c, v = op.upper_rows()
    a = g.arrays
    assert np.array_equal(rp, a["hh_rowptr"]) and np.array_equal(c, a["hh_cols"])
    assert np.max(np.abs(v - a["hh_vals"])) < 1e-12
    # the reference's symmetric row product on its rows vs both operator forms
    n = op.n
    X = np.sin(0.37 * np.arange(n)[:, None] + 1.3 * np.arange(3)[None, :])
    Yr = oracle.hh_times(n, a["hh_rowptr"], a["hh_cols"], a["hh_vals"], X)
    for flags in (gpu.OP_STORED, gpu.OP_FACTORED):
        Y = gpu.Operator.build(s, flags).apply(X)
        assert np.max(np.abs(Y - Yr)) <= 1e-12 * np.max(np.abs(Yr))


@pytest.mark.parametrize("name", __import__("conftest").TRUNC_CASES)
def test_truncated_rows(gpu, name):
    """Truncated bases (hw and occupation rejections): the GPU rows bit-exact against the
    oracle's on the same basis (with and without the 1e-14 drop), and against the rows
    the reference's XHH+JJ writes (structure exact, values within 1e-12); the row form
    regenerates the same products; the product form is refused."""
    from conftest import trunc_case
    xml, meta, a = trunc_case(name)
    s = gpu.System.from_xml(xml)
    basis = s.basis()
    o = oracle.System(xml)
    n = s.dim
    for keep in (False, True):
        op = gpu.Operator.build(s, gpu.OP_KEEP_SMALL if keep else gpu.OP_STORED)
        rp, c, v = op.upper_rows()
        orp, oc, ov = o.h_upper(basis=basis, drop_small=not keep)
        assert np.array_equal(rp, orp) and np.array_equal(c, oc)
        assert np.array_equal(v.view(np.uint64), ov.view(np.uint64))
        if keep and "hh_rowptr" in a:
            assert np.array_equal(rp, a["hh_rowptr"]) and np.array_equal(c, a["hh_cols"])
            assert np.max(np.abs(v - a["hh_vals"])) < 1e-12
        if keep and "sample_rows" in a:
            for i, r in enumerate(a["sample_rows"]):
                s0, s1 = a["sample_rowptr"][i], a["sample_rowptr"][i + 1]
                assert np.array_equal(c[rp[r]:rp[r + 1]], a["sample_cols"][s0:s1])
                assert np.max(np.abs(v[rp[r]:rp[r + 1]] - a["sample_vals"][s0:s1])) < 1e-12
        if not keep:
            X = np.random.default_rng(1).standard_normal((n, 5))
            ref = oracle.csb_times(n, orp, oc, ov, X)
            y = op.apply(X, gpu.PREC_FP64)
            assert np.max(np.abs(y - ref)) <= 1e-12 * np.abs(ref).max()
            vop = gpu.Operator.build(s, gpu.OP_VIRTUAL)
            assert np.max(np.abs(vop.apply(X, gpu.PREC_FP64) - y)) <= 1e-12 * np.abs(ref).max()
            vop.close()
        op.close()
    with pytest.raises(gpu.CosmoError):
        gpu.Operator.build(s, gpu.OP_FACTORED)


@pytest.mark.parametrize("name", ["psd2p2n_hw2", "sd3p3n_rej"])
def test_truncated_solve_and_labels(gpu, name):
    """Davidson on a truncated basis: the lowest eigenvalues of the oracle's dense H on the
    same basis within 1e-8, residuals <= 1e-8; J labels integral (hw and occupation
    rejections keep every level's occupation, so J is a good quantum number there)."""
    from conftest import trunc_case
    xml, meta, a = trunc_case(name)
    s = gpu.System.from_xml(xml)
    o = oracle.System(xml)
    rp, c, v = o.h_upper(basis=s.basis(), drop_small=True)
    w = np.linalg.eigvalsh(oracle.dense_from_upper(s.dim, rp, c, v))
    op = gpu.Operator.build(s)
    ev, de, rs, vec, st = gpu.davidson(op, 6, lim=40)
    assert np.max(np.abs(ev - w[:6])) < 1e-8 and rs.max() <= 1e-8
    jj, tt = s.labels(vec)
    lib = gpu.lib()
    for x in jj:
        J2 = lib.cosmo_twice_spin(float(x), 0)
        assert J2 >= 0 and abs(x - (J2 * (J2 + 2)) / 4.0) < 1e-6  # M = 0
    op.close()
