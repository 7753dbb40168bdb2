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
