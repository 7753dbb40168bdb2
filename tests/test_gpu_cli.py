"""The command-line surface: XHH+JJ -> .HH, davidson_vfile -> .EE, DEXHVJTcsb -> .EE + XML
states, on the golden systems (reference outputs in tests/golden)."""
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cosmo_amd", "bin")


def run(tool, *args, cwd):
    r = subprocess.run([os.path.join(BIN, tool)] + list(args), cwd=cwd, capture_output=True, text=True,
                       stdin=subprocess.DEVNULL, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    return r


def test_xhh_then_davidson_vfile(gpu, golden):
    g = golden("sd2p2n_s11")
    d = g.dir
    run("XHH+JJ", g.name + ".xml", cwd=d)
    hh = os.path.join(d, g.name + ".HH")
    rp, c, v = oracle.read_hh(hh)
    a = g.arrays
    assert np.array_equal(rp, a["hh_rowptr"]) and np.array_equal(c, a["hh_cols"])
    assert np.max(np.abs(v - a["hh_vals"])) < 1e-12
    root = ET.parse(os.path.join(d, g.name + ".xml")).getroot()
    m = root.find("system").find("matrix")
    assert m.get("name") == g.name + ".HH" and m.get("format") == "pp"
    run("davidson_vfile", g.name + ".HH", "-n", "5", cwd=d)
    E, dE, R, V = oracle.read_ee(hh + ".EE")
    assert len(E) == 5
    ref = np.array(g.meta["ref_evals"])[:5]
    assert np.max(np.abs(E - ref)) < 1e-8
    assert np.max(R) < 1e-6
    # warm start from the file just written
    run("davidson_vfile", g.name + ".HH", "-n", "5", "-f", g.name + ".HH.EE", cwd=d)


@pytest.mark.parametrize("limit", ["1000", "100"])
def test_dexhvjtcsb_states(gpu, golden, tmp_path, limit):
    g = golden("sd2p2n_s11")
    import shutil
    for ext in (".xml", ".int"):
        shutil.copy(os.path.join(g.dir, g.name + ext), str(tmp_path / (g.name + ext)))
    run("DEXHVJTcsb", g.name + ".xml", "-n", "6", "-l", limit, cwd=str(tmp_path))
    E, dE, R, V = oracle.read_ee(str(tmp_path / (g.name + ".HH.EE")))
    ref = np.array(g.meta["ref_evals"])
    assert len(E) == (640 if limit == "1000" else 6)
    assert np.max(np.abs(E[:6] - ref[:6])) < 1e-8
    root = ET.parse(str(tmp_path / (g.name + ".xml"))).getroot()
    states = root.find("system").findall("state")
    assert len(states) == 6
    Es = [float(s.get("E")) for s in states]
    assert Es == sorted(Es)
    for s in states:
        assert "/" not in s.get("J")  # even-A: integer spins
        assert s.find("wavefunction").get("file") == g.name + ".HH.EE"
    # the labels are the J of the eigenvectors: <J^2> = J(J+1) to numerical precision
    s = gpu.System.from_xml(str(tmp_path / (g.name + ".xml")))
    vec = np.stack(V[:6], axis=1)
    jj, tt = s.labels(vec)
    for x, st in zip(jj, sorted(states, key=lambda q: float(q.get("E")))):
        J = int(st.get("J"))
        assert abs(x - J * (J + 1)) < 1e-6  # M = 0: <J-J+> = J(J+1)


@pytest.mark.parametrize("name,twom,twotz", [("sd3p1n_m1_s7", 2, -2), ("fp3p2n_m1_s3", 1, -1)])
def test_labels_integrality(gpu, golden, name, twom, twotz):
    """<J-J+> = J(J+1) - M(M+1) with 2J of the parity of A; same for isospin"""
    g = golden(name)
    s = gpu.System.from_xml(g.xml)
    op = gpu.Operator.build(s)
    ev, de, rs, vec, st = gpu.davidson(op, 6, lim=40)
    jj, tt = s.labels(vec)
    lib = gpu.lib()
    for x in jj:
        J2 = lib.cosmo_twice_spin(float(x), twom)
        assert (J2 - twom) % 2 == 0 and J2 >= abs(twom)
        assert abs(x - (J2 * (J2 + 2) - twom * (twom + 2)) / 4.0) < 1e-6
    for x in tt:
        T2 = lib.cosmo_twice_spin(float(x), twotz)
        assert abs(x - (T2 * (T2 + 2) - twotz * (twotz + 2)) / 4.0) < 1e-6


LABEL_CASES = ["sd2p2n_s11_lab", "sd3p1n_m1_s7_labj", "fp2p2n_s5_labt"]


def _states(path):
    """<state> nodes of the system: attributes in file order, with their wavefunctions."""
    out = []
    for st in ET.parse(path).getroot().find("system").findall("state"):
        out.append(dict(attrs=[list(kv) for kv in st.attrib.items()],
                        wavefunctions=[[list(kv) for kv in w.attrib.items()] for w in st.findall("wavefunction")]))
    return out


def _label_case(name, tmp_path):
    import json
    import shutil
    from conftest import GOLDEN
    ref = json.load(open(os.path.join(GOLDEN, name + "_states.json")))
    shutil.copy(os.path.join(GOLDEN, name + "_in.xml"), str(tmp_path / (name + ".xml")))
    shutil.copy(os.path.join(GOLDEN, name + ".int"), str(tmp_path / (name + ".int")))
    return ref


@pytest.mark.parametrize("name", LABEL_CASES)
def test_xshljt_on_the_reference_ee(gpu, tmp_path, name):
    """XSHLJT on the .EE the reference's davidson_vfile wrote gives the <state> nodes the
    reference's XSHLJT wrote from it, attribute for attribute and in order: J and T labels,
    the system's J / T filters, the matrix's JJ / TT shifts taken off the energies, names
    J P(k), Ex, dE and the wavefunction references (tests/golden/*_states.json)."""
    import shutil
    from conftest import GOLDEN
    ref = _label_case(name, tmp_path)
    shutil.copy(os.path.join(GOLDEN, name + "_EE.bin"), str(tmp_path / (name + ".HH.EE")))
    run("XSHLJT", name + ".xml", "-n", str(ref["states_asked"]), cwd=str(tmp_path))
    assert _states(str(tmp_path / (name + ".xml"))) == ref["states"]
    # run again: the states are found (J, T and E within 0.001) and not duplicated
    run("XSHLJT", name + ".xml", "-n", str(ref["states_asked"]), cwd=str(tmp_path))
    assert _states(str(tmp_path / (name + ".xml"))) == ref["states"]


@pytest.mark.parametrize("name", LABEL_CASES)
def test_dexhvjtcsb_labels_as_the_reference(gpu, tmp_path, name):
    """DEXHVJTcsb (its own H and solve) on the same input: the reference XSHLJT's states in
    the same order with the same J, P, T and names, each pointing at the state of the same
    index in <name>.HH.EE; energies within 2e-4 MeV of its 6-digit values where the
    reference's Davidson converged that pair (residual <= 1e-6), Ex where the lowest state
    converged too.  (In the fp case the reference's coefficient test accepted pairs 2 and 7
    with residuals 0.84 and 0.12; this build's strict solve converges them.)"""
    from conftest import GOLDEN
    ref = _label_case(name, tmp_path)
    _, _, R, _ = oracle.read_ee(os.path.join(GOLDEN, name + "_EE.bin"))
    k = str(ref["states_asked"])
    run("DEXHVJTcsb", name + ".xml", "-n", k, "-b", k, cwd=str(tmp_path))
    got = _states(str(tmp_path / (name + ".xml")))
    assert len(got) == len(ref["states"])
    ids = [int(dict(b["wavefunctions"][0])["id"]) for b in ref["states"]]
    ground_ok = R[ids[0]] <= 1e-6
    for a, b, i in zip(got, ref["states"], ids):
        da, db = dict(a["attrs"]), dict(b["attrs"])
        for key in ("J", "P", "T", "name"):
            assert da[key] == db[key], (key, da, db)
        if R[i] <= 1e-6:
            assert abs(float(da["E"]) - float(db["E"])) < 2e-4, (da, db)
            if ground_ok:
                assert abs(float(da["Ex"]) - float(db["Ex"])) < 2e-4, (da, db)
        assert [dict(w)["id"] for w in a["wavefunctions"]] == [dict(w)["id"] for w in b["wavefunctions"]]
        assert dict(a["wavefunctions"][0])["file"] == name + ".HH.EE"
