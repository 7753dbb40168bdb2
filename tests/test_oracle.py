"""Pins the CPU oracle: against the reference's own outputs (golden fixtures made by the
reference programs), textbook values, brute-force second quantisation and invariants."""
import hashlib
import itertools
import math
import os

import numpy as np
import pytest

import oracle
from conftest import FULL_CASES, GOLDEN_CASES


def test_cg_textbook_values():
    assert oracle.cg(1, 1, 1, -1, 2, 0) == pytest.approx(1 / math.sqrt(2), abs=1e-15)
    assert oracle.cg(1, 1, 1, -1, 0, 0) == pytest.approx(1 / math.sqrt(2), abs=1e-15)
    assert oracle.cg(1, -1, 1, 1, 0, 0) == pytest.approx(-1 / math.sqrt(2), abs=1e-15)
    assert oracle.cg(2, 2, 2, -2, 0, 0) == pytest.approx(1 / math.sqrt(3), abs=1e-15)
    assert oracle.cg(2, 0, 2, 0, 0, 0) == pytest.approx(-1 / math.sqrt(3), abs=1e-15)
    assert oracle.cg(2, 2, 1, -1, 1, 1) == pytest.approx(math.sqrt(2 / 3), abs=1e-15)
    assert oracle.cg(3, 1, 1, 1, 4, 2) == pytest.approx(math.sqrt(3) / 2, abs=1e-15)
    assert oracle.cg(1, 1, 1, 1, 0, 2) == 0.0


@pytest.mark.parametrize("j1,j2", [(7, 5), (5, 5), (3, 1), (7, 7)])
def test_cg_orthogonality(j1, j2):
    for M in range(-(j1 + j2), j1 + j2 + 1, 2):
        Js = [J for J in range(abs(j1 - j2), j1 + j2 + 1, 2) if abs(M) <= J]
        ms = [(m1, M - m1) for m1 in range(-j1, j1 + 1, 2) if abs(M - m1) <= j2]
        C = np.array([[oracle.cg(j1, m1, j2, m2, J, M) for (m1, m2) in ms] for J in Js])
        assert np.allclose(C @ C.T, np.eye(len(Js)), atol=1e-14)


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_oracle_basis_and_tensors_vs_reference(golden, name):
    g = golden(name)
    o = oracle.System(g.xml)
    b = o.basis()
    raw = np.array([b.shape[1]], "<i4").tobytes() + np.array([b.shape[0]], "<u8").tobytes() + b.tobytes()
    assert hashlib.sha256(raw).hexdigest() == g.meta["mbs_sha256"]
    assert np.array_equal(o.vspx_keys, g.arrays["vspx_keys"])
    assert np.max(np.abs(o.vspx_vals - g.arrays["vspx_vals"])) < 5e-14
    assert np.array_equal(o.esp_vals, g.arrays["esp_vals"])


@pytest.mark.parametrize("name", FULL_CASES)
def test_oracle_rows_vs_reference_hh(golden, name):
    g = golden(name)
    rp, c, v = oracle.System(g.xml).h_upper(drop_small=False)
    a = g.arrays
    assert np.array_equal(rp, a["hh_rowptr"]) and np.array_equal(c, a["hh_cols"])
    assert np.max(np.abs(v - a["hh_vals"])) < 1e-12


@pytest.mark.parametrize("name", FULL_CASES)
def test_oracle_csb_product_bit_exact_vs_reference(golden, name):
    """oracle csb_times on the reference's own H values reproduces SparseMatrixCSB::TimesMatrix
    (run by oracle/_ref/ref_harness) bit for bit, including its 1e-14 filter"""
    g = golden(name)
    a = g.arrays
    rp, c, v = a["hh_rowptr"], a["hh_cols"], a["hh_vals"]
    n = g.meta["n"]
    keep = (np.abs(v) >= 1e-14)
    cnt = np.array([keep[rp[i]:rp[i + 1]].sum() for i in range(n)])
    rpk = np.concatenate([[0], np.cumsum(cnt)])
    assert rpk[-1] == g.meta["nnz_upper_csb"]
    Y = oracle.csb_times(n, rpk, c[keep], v[keep], a["spmm_x"])
    assert np.array_equal(Y.view(np.uint64), a["spmm_y"].view(np.uint64))


def _brute_h(o, basis):
    """H from explicit fermion operators on bit strings: sum over forward terms of
    v (T + T^dagger), T = a+_{b2} a+_{b3} a_{b1} a_{b0}; diagonal keys and one-body
    number terms once.  Independent of the oracle's list walking."""
    def ann(state, i):
        if not (state >> i) & 1:
            return None, 0
        sgn = bin(state & ((1 << i) - 1)).count("1") & 1
        return state & ~(1 << i), sgn

    def cre(state, i):
        if (state >> i) & 1:
            return None, 0
        sgn = bin(state & ((1 << i) - 1)).count("1") & 1
        return state | (1 << i), sgn

    def apply(ops, state):  # ops applied right to left: list of (kind, orbital)
        s, ph = state, 0
        for kind, i in reversed(ops):
            s, p = (ann if kind == "a" else cre)(s, i)
            if s is None:
                return None, 0
            ph ^= p
        return s, ph

    idx = {}
    masks = []
    for r, row in enumerate(basis):
        m = 0
        for x in row:
            m |= 1 << int(x)
        idx[m] = r
        masks.append(m)
    n = len(basis)
    H = np.zeros((n, n))
    for (i, j), v in zip(o.esp_keys, o.esp_vals):
        for r, m in enumerate(masks):
            s, ph = apply([("c", int(j)), ("a", int(i))], m)
            if s is not None and s in idx:
                H[idx[s], r] += -v if ph else v
    for k, v in zip(o.vspx_keys, o.vspx_vals):
        b0, b1, b2, b3 = (int(x) for x in k)
        T = [("c", b2), ("c", b3), ("a", b1), ("a", b0)]
        Td = [("c", b0), ("c", b1), ("a", b3), ("a", b2)]
        for r, m in enumerate(masks):
            s, ph = apply(T, m)
            if s is not None and s in idx:
                H[idx[s], r] += -v if ph else v
            if (b0, b1) != (b2, b3):
                s, ph = apply(Td, m)
                if s is not None and s in idx:
                    H[idx[s], r] += -v if ph else v
    return H


@pytest.mark.parametrize("shell,nv,zv,twom", [("sd", 1, 1, 0), ("sd", 2, 1, 1), ("fp", 1, 2, 3)])
def test_oracle_h_vs_brute_force_second_quantisation(tmp_path, shell, nv, zv, twom):
    from cosmo_amd import synth
    par = "-" if (shell == "fp" and (nv + zv) % 2) else "+"
    xml = synth.write_system(str(tmp_path), shell, "bf", nv, zv, twom, par, 9)
    o = oracle.System(xml)
    b = o.basis()
    rp, c, v = o.h_upper(b, drop_small=False)
    H = oracle.dense_from_upper(len(b), rp, c, v)
    Hb = _brute_h(o, b)
    assert np.max(np.abs(H - Hb)) < 1e-13


def test_rotational_invariance_of_spectrum(tmp_path):
    """[H, J^2] = 0: every level of the M=1 space appears in the M=0 space"""
    from cosmo_amd import synth
    ev = {}
    for twom in (0, 2):
        xml = synth.write_system(str(tmp_path), "sd", "ri%d" % twom, 1, 2, twom, "+", 4)
        o = oracle.System(xml)
        b = o.basis()
        rp, c, v = o.h_upper(b)
        ev[twom] = np.linalg.eigvalsh(oracle.dense_from_upper(len(b), rp, c, v))
    for e in ev[2]:
        assert np.min(np.abs(ev[0] - e)) < 1e-9


REF_USDB = "/root/reference/interactions/sd/usdb.int"


@pytest.mark.skipif(not os.path.exists(REF_USDB), reason="reference interaction database not present")
def test_known_answer_24mg_usdb(tmp_path):
    """EXAMPLE-24Mg-usdb.md:29-38 (reference's published levels; state 5 excluded: the
    reference's own Davidson leaves it unconverged)"""
    import shutil
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    shutil.copy(REF_USDB, str(tmp_path / "usdb.int"))
    xml = tmp_path / "mg24.xml"
    xml.write_text('<cosmo><valencespace name="sd"><core name="16O" N="8" Z="8"/>'
                   '<orbital name="0d3" n="0" l="2" j="3/2" type="pn" index="1"/>'
                   '<orbital name="0d5" n="0" l="2" j="5/2" type="pn" index="2"/>'
                   '<orbital name="1s1" n="1" l="0" j="1/2" type="pn" index="3"/></valencespace>'
                   '<hamiltonian name="usdb"><file><name>usdb.int</name></file></hamiltonian>'
                   '<system name="mg24" A="24" Av="8" Nv="4" Zv="4" Tz="0" Jz="0" P="+"/></cosmo>')
    o = oracle.System(str(xml))
    b = o.basis()
    assert len(b) == 28503
    rp, c, v = o.h_upper(b)
    n = len(b)
    rows = np.repeat(np.arange(n), np.diff(rp))
    U = sp.csr_matrix((v, (rows, c)), shape=(n, n))
    H = U + sp.triu(U, 1).T
    w = np.sort(sla.eigsh(H, k=12, which="SA", tol=1e-10)[0])[:10]
    pub = [-87.1044, -85.6021, -82.9883, -82.732, -82.0294, -81.2219, -79.7662, -79.6227, -79.3076, -79.2863]
    for i, e in enumerate(pub):
        if i == 4:
            continue
        assert abs(w[i] - e) < 6e-5, (i, w[i], e)


def _filtered_upper(g):
    """The reference's own .HH rows with the CSB builder's |h| < 1e-14 drop applied
    (include/sm/PPMatrix.cxx:621)."""
    n = g.meta["n"]
    rp, c, v = g.arrays["hh_rowptr"], g.arrays["hh_cols"], g.arrays["hh_vals"]
    keep = np.abs(v) >= 1e-14
    rows = np.repeat(np.arange(n), np.diff(rp))
    rp2 = np.zeros(n + 1, np.int64)
    np.add.at(rp2, rows[keep] + 1, 1)
    return np.cumsum(rp2), c[keep], v[keep]


@pytest.mark.parametrize("name", FULL_CASES)
def test_oracle_csb_file_is_the_reference_file(golden, name, tmp_path):
    """The oracle's CSB layout applied to the reference's own rows reproduces, byte for
    byte, the file SparseMatrixCSB::Write made (its SHA-256 and size); reading it back
    returns the same rows."""
    g = golden(name)
    rp, c, v = _filtered_upper(g)
    assert rp[-1] == g.meta["nnz_upper_csb"]
    b = oracle.csb_bytes(g.meta["n"], rp, c, v)
    assert len(b) == g.meta["csb_bytes"]
    assert hashlib.sha256(b).hexdigest() == g.meta["csb_sha256"]
    p = str(tmp_path / "h.csb")
    open(p, "wb").write(b)
    hdr, blocks, isha = oracle.read_csb(p)
    assert isha == g.meta["csb_index_sha256"]
    assert hdr["br"] == g.meta["csb_block_rows"] and hdr["nbr"] == g.meta["csb_num_block_rows"]
    rp2, c2, v2 = oracle.csb_to_upper(hdr, blocks)
    assert np.array_equal(rp2, rp) and np.array_equal(c2, c) and np.array_equal(v2, v)


@pytest.mark.parametrize("name", FULL_CASES)
def test_oracle_h_gives_the_reference_csb_structure(golden, name):
    """The oracle's own H, laid out as CSB, has the reference file's header, block flags,
    counts and indices (its values are pinned against the reference's .HH rows, to 1e-12)."""
    g = golden(name)
    o = oracle.System(g.xml)
    rp, c, v = o.h_upper(drop_small=True)
    b = oracle.csb_bytes(g.meta["n"], rp, c, v)
    import tempfile
    with tempfile.NamedTemporaryFile(suffix=".csb") as f:
        f.write(b)
        f.flush()
        _, _, isha = oracle.read_csb(f.name)
    assert isha == g.meta["csb_index_sha256"]


def _sin_x(n, m=3):
    """The deterministic input of the reference runs: x[v][i] = sin(0.37 i + 1.3 v)."""
    i = np.arange(n, dtype=np.float64)[:, None]
    v = np.arange(m, dtype=np.float64)[None, :]
    return np.sin(0.37 * i + 1.3 * v)


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_oracle_hh_product_bit_exact_vs_reference(golden, name):
    """oracle hh_times (orc_hh_times) reproduces SparseMatrix<double>::TimesMatrix with type 1
    (include/tav/sparse.cxx:273-291, run by oracle/_ref/ref_harness hhtimes on the .HH the
    reference's XHH+JJ wrote) bit for bit: on the reference's own rows where the fixture
    holds them, on the oracle's rows (their SHA-256 pins every element) otherwise."""
    g = golden(name)
    n = g.meta["n"]
    a = g.arrays
    X = _sin_x(n)
    if "spmm_x" in a:  # the same doubles as the reference's input
        assert np.array_equal(X.view(np.uint64), a["spmm_x"].view(np.uint64))
    if "hh_y" in a:
        Y = oracle.hh_times(n, a["hh_rowptr"], a["hh_cols"], a["hh_vals"], X)
        assert np.array_equal(Y.view(np.uint64), a["hh_y"].view(np.uint64))
        assert hashlib.sha256(np.ascontiguousarray(Y.T).tobytes()).hexdigest() == g.meta["hh_y_sha256"]
    rp, c, v = oracle.System(g.xml).h_upper(drop_small=False)
    Y = oracle.hh_times(n, rp, c, v, X)
    ref, got = (a["hh_y"], Y) if "hh_y" in a else (a["hh_y_sample"], Y[a["sample_rows"]])
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))  # the oracle's values: CG noise


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_oracle_csb_port_vs_reference(golden, name):
    """The CPU-baseline port of the CSB product (orc_csb_build / orc_csb_apply, row-major
    vectors) returns SparseMatrixCSB::TimesMatrix's output bit for bit on the reference's
    rows (full cases), and agrees with it to 1e-12 relative on the oracle's rows."""
    g = golden(name)
    n = g.meta["n"]
    a = g.arrays
    X = _sin_x(n)
    if "hh_rowptr" in a:
        rp, c, v = _filtered_upper(g)
        Y = oracle.csb_port_times(n, rp, c, v, X)
        assert np.array_equal(Y.view(np.uint64), a["spmm_y"].view(np.uint64))
    rp, c, v = oracle.System(g.xml).h_upper(drop_small=True)
    Y = oracle.csb_port_times(n, rp, c, v, X)
    ref = a["spmm_y"] if "spmm_y" in a else a["spmm_y_sample"]
    got = Y if "spmm_y" in a else Y[a["sample_rows"]]
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    # and the port equals the pinned csb_times on the same rows, bit for bit
    Yc = oracle.csb_times(n, rp, c, v, X)
    assert np.array_equal(Y.view(np.uint64), np.ascontiguousarray(Yc).view(np.uint64))


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_oracle_mbs_reader_on_the_reference_bytes(golden, name, tmp_path):
    """read_mbs on a file whose SHA-256 is the reference XsysmbsOMP output's returns the
    oracle's basis (the file is written by the library's .mbs writer, which needs no GPU)."""
    import cosmo_amd
    g = golden(name)
    p = str(tmp_path / "b.mbs")
    cosmo_amd.System.from_xml(g.xml).write_mbs(p)
    assert hashlib.sha256(open(p, "rb").read()).hexdigest() == g.meta["mbs_sha256"]
    assert np.array_equal(oracle.read_mbs(p), oracle.System(g.xml).basis())


def test_oracle_ee_reader_round_trip(tmp_path):
    """read_ee against the library's .EE writer: counts, values, residuals and vectors bit
    for bit; a file cut short fails to parse."""
    import cosmo_amd
    rng = np.random.default_rng(4)
    n, k = 37, 3
    E, dE, R = rng.standard_normal(k), rng.standard_normal(k) * 1e-9, rng.random(k) * 1e-7
    V = rng.standard_normal((k, n))
    p = str(tmp_path / "x.EE")
    cosmo_amd.write_ee(p, E, dE, R, V.T)  # vectors as columns, n x k
    e2, d2, r2, v2 = oracle.read_ee(p)
    assert np.array_equal(e2, E) and np.array_equal(d2, dE) and np.array_equal(r2, R)
    assert all(np.array_equal(a, b) for a, b in zip(v2, V))
    assert os.path.getsize(p) == 4 + k * (32 + 8 * n)
    open(p, "r+b").truncate(4 + 32 + 8 * n + 16)
    with pytest.raises(Exception):
        oracle.read_ee(p)


@pytest.mark.parametrize("name", __import__("conftest").TRUNC_CASES)
def test_oracle_rows_on_truncated_basis_vs_reference_hh(name):
    """The oracle's row generator on a truncated basis (the system's, whose .mbs is the
    reference's byte for byte) gives the rows the reference's XHH+JJ writes for it: hops
    leaving the basis dropped, structure exact, values within 1e-12."""
    import cosmo_amd
    from conftest import trunc_case
    xml, meta, a = trunc_case(name)
    basis = cosmo_amd.System.from_xml(xml).basis()
    o = oracle.System(xml)
    if "hh_rowptr" in a:
        rp, c, v = o.h_upper(basis=basis, drop_small=False)
        assert np.array_equal(rp, a["hh_rowptr"]) and np.array_equal(c, a["hh_cols"])
        assert np.max(np.abs(v - a["hh_vals"])) < 1e-12
    else:
        for i, r in enumerate(a["sample_rows"]):
            rp, c, v = o.h_upper(basis=basis, drop_small=False, r0=int(r), r1=int(r) + 1)
            s0, s1 = a["sample_rowptr"][i], a["sample_rowptr"][i + 1]
            assert np.array_equal(c, a["sample_cols"][s0:s1])
            assert np.max(np.abs(v - a["sample_vals"][s0:s1])) < 1e-12
