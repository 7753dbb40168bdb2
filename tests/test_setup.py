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
