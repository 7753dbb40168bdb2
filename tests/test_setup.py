"""Host-side setup of the product (XML, Oxbash reader, J->m, basis) against the oracle and
against the reference's own outputs (golden fixtures)."""
import hashlib
import os

import numpy as np
import pytest

import cosmo_amd
import oracle
from conftest import FULL_CASES, GOLDEN_CASES


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_basis_matches_reference_mbs(golden, name):
    g = golden(name)
    s = cosmo_amd.System.from_xml(g.xml)
    assert s.dim == g.meta["n"]
    b = s.basis()
    z = b.astype("<u2").tobytes()
    raw = np.array([b.shape[1]], "<i4").tobytes() + np.array([b.shape[0]], "<u8").tobytes() + z
    assert hashlib.sha256(raw).hexdigest() == g.meta["mbs_sha256"]


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_mscheme_tensors_match_reference(golden, name):
    g = golden(name)
    s = cosmo_amd.System.from_xml(g.xml)
    ek, ev, vk, vv = s.tensors()
    assert np.array_equal(ek, g.arrays["esp_keys"])
    assert np.array_equal(ev, g.arrays["esp_vals"])  # SPEs are exact (float parse, no CG)
    assert np.array_equal(vk, g.arrays["vspx_keys"])  # keys exact
    # values: CG by closed form vs the reference's recurrence
    assert np.max(np.abs(vv - g.arrays["vspx_vals"])) < 5e-14


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_product_setup_equals_oracle(golden, name):
    g = golden(name)
    s = cosmo_amd.System.from_xml(g.xml)
    o = oracle.System(g.xml)
    assert np.array_equal(s.basis(), o.basis())
    ek, ev, vk, vv = s.tensors()
    assert np.array_equal(vk, o.vspx_keys) and np.array_equal(vv, o.vspx_vals)


def test_closed_form_dimensions(tmp_path):
    from cosmo_amd import synth
    for shell, nv, zv, n in [("sd", 4, 4, 28503), ("sd", 6, 6, 93710), ("fp", 4, 4, 1963461)]:
        xml = synth.write_system(str(tmp_path), shell, "d%s%d" % (shell, nv), nv, zv, 0, "+", 1)
        assert cosmo_amd.System.from_xml(xml).dim == n


def test_brute_force_basis_tiny(tmp_path):
    """every sorted occupation list with the right counts, M and parity, in lex order"""
    import itertools
    from cosmo_amd import synth
    xml = synth.write_system(str(tmp_path), "sd", "t", 2, 1, 1, "+", 2)
    s = cosmo_amd.System.from_xml(xml)
    o = oracle.System(xml)
    m2 = [x["m2"] for x in o.sps]
    want = []
    for c in itertools.combinations(range(24), 3):
        if sum(1 for x in c if x < 12) == 2 and sum(m2[x] for x in c) == 1:
            want.append(c)
    assert np.array_equal(s.basis(), np.array(sorted(want), dtype=np.uint16))


def test_mbs_roundtrip_and_truncation(golden, tmp_path):
    g = golden("sd2p2n_s11")
    s = cosmo_amd.System.from_xml(g.xml)
    p = str(tmp_path / "x.mbs")
    s.write_mbs(p)
    s.check_mbs(p)
    data = open(p, "rb").read()
    open(p, "wb").write(data[:-2])
    with pytest.raises(cosmo_amd.CosmoError):
        s.check_mbs(p)


def test_missing_interaction_file(tmp_path):
    from cosmo_amd import synth
    xml = synth.write_system(str(tmp_path), "sd", "t", 1, 1, 0, "+", 2)
    os.remove(str(tmp_path / "t.int"))
    with pytest.raises(cosmo_amd.CosmoError) as e:
        cosmo_amd.System.from_xml(xml)
    assert e.value.code == cosmo_amd.EIO


MATRIX_CASES = ["sd2p2n_s11_jjtt", "fp2p2n_s5_jj", "sd3p1n_m1_s7_tt"]


@pytest.mark.parametrize("name", MATRIX_CASES)
def test_jj_tt_tensors_match_reference(golden, name):
    """<matrix JJ=".." TT=".."> adds scale * J-J+ / T-T+ to ESP and VSPX as the reference's
    xReadSPInt does: ESP bit-exact and the VSPX keys exact against its dump, values within
    the CG noise; the reference's two-body keys that annihilate or create one orbital
    twice (they never act on a state) are the only ones this build leaves out."""
    g = golden(name)
    s = cosmo_amd.System.from_xml(g.xml)
    ek, ev, vk, vv = s.tensors()
    a = g.arrays
    assert np.array_equal(ek, a["esp_keys"]) and np.array_equal(ev, a["esp_vals"])
    rk, rv = a["vspx_keys"], a["vspx_vals"]
    acts = (rk[:, 0] != rk[:, 1]) & (rk[:, 2] != rk[:, 3])
    assert np.array_equal(vk, rk[acts])
    assert np.max(np.abs(vv - rv[acts])) < 5e-14
    base = golden(g.meta["base"])
    # the add-on changes the operator: the one-body part or the two-body values differ
    assert not (np.array_equal(ev, base.arrays["esp_vals"]) and np.array_equal(vk, base.arrays["vspx_keys"])
                and np.max(np.abs(vv - base.arrays["vspx_vals"])) < 1e-12)


def test_jj_shifts_only_by_j_squared(tmp_path):
    """H + c J-J+ has the eigenvectors of H, each level moved by c (J(J+1) - M(M+1)): at M = 0
    the shifts are c J(J+1) with J integral, so every shifted level minus an unshifted one
    is a multiple of c that is J(J+1) for some J (the oracle's dense H on both tensors)."""
    from cosmo_amd import synth
    ev = {}
    for tag, attrs in (("a", None), ("b", {"JJ": "0.75"})):
        xml = synth.write_system(str(tmp_path), "sd", "jj" + tag, 1, 2, 1, "+", 5, matrix_attrs=attrs)
        s = cosmo_amd.System.from_xml(xml)
        ek, evv, vk, vv = s.tensors()
        o = oracle.System(xml)
        o.esp_keys, o.esp_vals, o.vspx_keys, o.vspx_vals = (np.ascontiguousarray(ek), np.ascontiguousarray(evv),
                                                            np.ascontiguousarray(vk), np.ascontiguousarray(vv))
        b = o.basis()
        rp, c, v = o.h_upper(b, drop_small=False)
        H = oracle.dense_from_upper(len(b), rp, c, v)
        ev[tag] = (np.linalg.eigh(H), H)
    (w0, V0), H0 = ev["a"]
    (w1, V1), H1 = ev["b"]
    # the added operator in the eigenbasis of H: diagonal, c (J(J+1) - M(M+1)) per level
    D = V0.T @ (H1 - H0) @ V0
    allowed = [0.75 * (j2 * (j2 + 2) - 3) / 4.0 for j2 in range(1, 16, 2)]  # 2J odd, 2M = 1
    for d in np.diag(D):
        assert min(abs(d - a) for a in allowed) < 1e-9, d
    assert np.max(np.abs(D - np.diag(np.diag(D)))) < 1e-9
    assert np.max(np.abs(H0 @ H1 - H1 @ H0)) < 1e-12  # [H, J-J+] = 0


@pytest.mark.parametrize("where,snippet", [
    ("matrix", ' Lawson="1.0"'), ("matrix", ' Coulomb="1.44"'), ("matrix", ' Delta="0.5"'),
    ("ham", '<Lawson>2</Lawson>'), ("ham", '<tagfile fileName="x"/>'), ("ham", '<jispfile name="x"/>')])
def test_unsupported_hamiltonian_terms_are_refused(tmp_path, where, snippet):
    """Terms of xReadSPInt this build does not construct fail with COSMO_EUNSUP instead of
    being ignored (the H would silently differ from the reference's)."""
    from cosmo_amd import synth
    xml = synth.write_system(str(tmp_path), "sd", "u", 1, 1, 0, "+", 2, matrix_attrs={"JJ": "0"})
    text = open(xml).read()
    if where == "matrix":
        text = text.replace('format="pp" JJ="0"', 'format="pp"' + snippet)
    else:
        text = text.replace("</hamiltonian>", snippet + "</hamiltonian>")
    open(xml, "w").write(text)
    with pytest.raises(cosmo_amd.CosmoError) as e:
        cosmo_amd.System.from_xml(xml)
    assert e.value.code == cosmo_amd.EUNSUP


def test_av_and_tz_give_the_valence_numbers(tmp_path):
    """Without Nv / Zv the system's Av and Tz (physical (N - Z) / 2, as the reference writes
    it) fix the valence numbers (include/xsm/xManyBodyStates.cxx:47-55): the same basis as
    with Nv and Zv; a Tz of the wrong parity for Av is refused."""
    import re
    from cosmo_amd import synth
    for (nv, zv, tz) in ((1, 3, "-1"), (2, 3, "-1/2"), (4, 2, "1")):
        xml = synth.write_system(str(tmp_path), "sd", "s%d%d" % (nv, zv), nv, zv, (nv + zv) % 2, "+", 3)
        n_ref = cosmo_amd.System.from_xml(xml).dim
        txt = open(xml).read()
        assert ('Tz="%s"' % tz) in txt
        open(xml, "w").write(re.sub(r' Nv="\d+" Zv="\d+"', "", txt))
        assert cosmo_amd.System.from_xml(xml).dim == n_ref
    open(xml, "w").write(open(xml).read().replace('Tz="1"', 'Tz="1/2"'))
    with pytest.raises(cosmo_amd.CosmoError):
        cosmo_amd.System.from_xml(xml)


def test_twice_spin_from_jmjp():
    """2J from <J-J+> = J(J+1) - M(M+1) for integer and half-integer spins; -1 when J + M is
    not an integer to within 0.1 (include/sm/QN.cxx:40-55)."""
    lib = cosmo_amd.lib()
    for J2 in range(0, 13):
        for M2 in range(-J2, J2 + 1, 2):
            x = (J2 * (J2 + 2) - M2 * (M2 + 2)) / 4.0
            assert lib.cosmo_twice_spin(x + 0.03, M2) == J2
    assert lib.cosmo_twice_spin(1.0, 0) == -1   # J = 0.618
    assert lib.cosmo_twice_spin(-5.0, 0) == -1  # no real spin


@pytest.mark.parametrize("name", __import__("conftest").TRUNC_CASES)
def test_truncated_basis_matches_reference_mbs(name, tmp_path):
    """The hw and occupation rejections of XsysmbsOMP (xmlshell/xmbs/XsysmbsOMP.cpp:212-303):
    the basis the system builds is the reference's .mbs byte for byte; that file passes the
    check, and the same system without its truncation refuses it."""
    import re
    from conftest import trunc_case
    xml, meta, _ = trunc_case(name)
    s = cosmo_amd.System.from_xml(xml)
    assert s.dim == meta["n"]
    p = str(tmp_path / "t.mbs")
    s.write_mbs(p)
    assert hashlib.sha256(open(p, "rb").read()).hexdigest() == meta["mbs_sha256"]
    s.check_mbs(p)
    txt = open(xml).read()
    full = re.sub(r'\s*<hwrejection[^>]*/>|\s*<rejection.*?</rejection>', "", txt, flags=re.S)
    full = full.replace(' hwmax="2"', "")
    q = str(tmp_path / "full.xml")
    open(q, "w").write(full)
    import shutil
    shutil.copy(xml[:-4] + ".int", str(tmp_path / (name + ".int")))
    f = cosmo_amd.System.from_xml(q)
    assert f.dim > s.dim
    with pytest.raises(cosmo_amd.CosmoError):
        f.check_mbs(p)


def test_truncation_input_errors(tmp_path):
    """Malformed truncations are refused: a rejection without NMAX, an orbital index of a
    level the species does not have, an hwrejection without HWMIN."""
    from cosmo_amd import synth
    bad = ['<rejection NMIN="0">\n<orbital index="1"/>\n</rejection>',
           '<rejection NMIN="0" NMAX="2">\n<orbital index="9" type="pn"/>\n</rejection>',
           '<hwrejection HWMAX="2"/>']
    for i, kids in enumerate(bad):
        xml = synth.write_system(str(tmp_path), "psd", "b%d" % i, 2, 2, 0, "+", 1, system_children=kids)
        with pytest.raises(cosmo_amd.CosmoError):
            cosmo_amd.System.from_xml(xml)


def test_empty_basis_is_refused_before_any_device_work(tmp_path):
    """An odd number of valence nucleons at 2M = 0 has no M-scheme state: the basis is
    empty and building an operator on it fails with COSMO_EINVAL (a clear message, before
    any device call)."""
    import cosmo_amd as ca
    from cosmo_amd import synth
    s = ca.System.from_xml(synth.write_system(str(tmp_path), "sd", "odd", 2, 3, 0, "+", 1))
    assert s.dim == 0
    with pytest.raises(ca.CosmoError) as e:
        ca.Operator.build(s)
    assert e.value.code == ca.EINVAL and "empty basis" in str(e.value)
