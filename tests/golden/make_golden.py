"""Writes the golden fixtures of tests/golden from the reference's own programs.

Run only where the CoSMo source tree is present at /root/reference (the committed
fixtures are what the tests read elsewhere): ``make -C oracle ref`` builds XsysmbsOMP, XHH+JJ, davidson_vfile, XSHLJT and ref_harness
into oracle/_ref from those sources, then this script runs them on seeded synthetic
systems and stores inputs and outputs as small data files (npz / json / text).
``python make_golden.py csb|hh|traj|matrix|labels|trunc`` only adds the CSB file keys, the
.HH row product, the Davidson counter runs, the <matrix> JJ/TT cases, the state labelling
(XSHLJT) cases or the truncated-basis cases.
"""
import hashlib
import json
import os
import shutil
import struct
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402
from cosmo_amd import synth  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
OUT = os.path.dirname(os.path.abspath(__file__))

# name, shell, Nv, Zv, 2Jz, parity, seed, full (commit .HH rows) , davidson block
CASES = [
    ("sd2p2n_s11", "sd", 2, 2, 0, "+", 11, True, 10),
    ("sd3p1n_m1_s7", "sd", 1, 3, 2, "+", 7, True, 6),
    ("fp2p2n_s5", "fp", 2, 2, 0, "+", 5, True, 10),
    ("fp3p2n_m1_s3", "fp", 2, 3, 1, "-", 3, False, 8),
    ("sd4p4n_s0", "sd", 4, 4, 0, "+", 0, False, 10),
]


def run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, stdin=subprocess.DEVNULL, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("%s failed: %s" % (cmd, r.stderr[-2000:]))
    return r.stdout


def read_dump(path):
    data = open(path, "rb").read()
    off = 0
    out = []
    for _ in range(2):
        r, = struct.unpack_from("<I", data, off)
        c, = struct.unpack_from("<Q", data, off + 4)
        off += 12
        keys = np.zeros((c, r), np.uint16)
        vals = np.zeros(c)
        for i in range(c):
            keys[i] = struct.unpack_from("<%dH" % r, data, off)
            off += 2 * r
            vals[i], = struct.unpack_from("<d", data, off)
            off += 8
        out.append((keys, vals))
    return out


def last_json(text):
    return json.loads([ln for ln in text.splitlines() if ln.startswith("{")][-1])


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def csb_meta(d, name):
    """The reference's CSB file of the system in d: its size, SHA-256, and the SHA-256 of
    everything but the values (header, block flags, counts and indices)."""
    run([os.path.join(REF, "ref_harness"), "csbwrite", name + ".xml", "ref.csb"], d)
    path = os.path.join(d, "ref.csb")
    hdr, blocks, idx_sha = oracle.read_csb(path)
    return dict(csb_bytes=os.path.getsize(path), csb_sha256=sha(path), csb_index_sha256=idx_sha,
                csb_block_rows=hdr["br"], csb_num_block_rows=hdr["nbr"])


def csb_only():
    """Adds the CSB file keys to the existing JSON fixtures (the rest is left as is)."""
    for (name, shell, nv, zv, twojz, par, seed, full, bloc) in CASES:
        d = tempfile.mkdtemp()
        try:
            synth.write_system(d, shell, name, nv, zv, twojz, par, seed)
            run([os.path.join(REF, "XsysmbsOMP"), name + ".xml"], d)
            jp = os.path.join(OUT, name + ".json")
            meta = json.load(open(jp))
            meta.update(csb_meta(d, name))
            with open(jp, "w") as f:
                json.dump(meta, f, indent=1)
            print(name, meta["csb_bytes"], meta["csb_index_sha256"][:16])
        finally:
            shutil.rmtree(d)


def main():
    for (name, shell, nv, zv, twojz, par, seed, full, bloc) in CASES:
        d = tempfile.mkdtemp()
        try:
            xml = synth.write_system(d, shell, name, nv, zv, twojz, par, seed)
            shutil.copy(os.path.join(d, name + ".int"), os.path.join(OUT, name + ".int"))
            shutil.copy(xml, os.path.join(OUT, name + ".xml"))
            run([os.path.join(REF, "XsysmbsOMP"), name + ".xml"], d)
            run([os.path.join(REF, "ref_harness"), "dumpint", name + ".xml", "dump.bin"], d)
            (ek, ev), (vk, vv) = read_dump(os.path.join(d, "dump.bin"))
            run([os.path.join(REF, "XHH+JJ"), name + ".xml"], d)
            info = last_json(run([os.path.join(REF, "ref_harness"), "spmm", name + ".xml", "3", "1", "spmm.bin"], d))
            raw = open(os.path.join(d, "spmm.bin"), "rb").read()
            n, m = struct.unpack_from("<QQ", raw, 0)
            X = np.frombuffer(raw, np.float64, n * m, 16).reshape(m, n).T
            Y = np.frombuffer(raw, np.float64, n * m, 16 + 8 * n * m).reshape(m, n).T
            dinfo = last_json(run([os.path.join(REF, "ref_harness"), "csb", name + ".xml", "ref.EE", str(bloc), "0",
                                    "0"], d))
            E, dE, R, V = oracle.read_ee(os.path.join(d, "ref.EE"))
            mbs = oracle.read_mbs(os.path.join(d, name + ".mbs"))
            meta = dict(name=name, shell=shell, nv=nv, zv=zv, twojz=twojz, parity=par, seed=seed, n=int(n),
                        nnz_upper_csb=int(info["nnz_upper"]), block_rows=int(info["block_rows"]),
                        mbs_sha256=sha(os.path.join(d, name + ".mbs")), hh_sha256=sha(os.path.join(d, name + ".HH")),
                        davidson_block=bloc, ref_evals=E.tolist(), ref_dE=dE.tolist(), ref_resid=R.tolist())
            rp, cols, vals = oracle.read_hh(os.path.join(d, name + ".HH"))
            meta["hh_nnz"] = int(rp[-1])
            meta.update(csb_meta(d, name))
            arrays = dict(esp_keys=ek, esp_vals=ev, vspx_keys=vk, vspx_vals=vv)
            if full:
                arrays.update(mbs=mbs, hh_rowptr=rp, hh_cols=cols, hh_vals=vals, spmm_x=X, spmm_y=Y)
            else:
                rng = np.random.default_rng(1)
                rows = np.sort(rng.choice(n, 16, replace=False))
                arrays.update(sample_rows=rows,
                              sample_rowptr=np.concatenate([[0], np.cumsum(rp[rows + 1] - rp[rows])]),
                              sample_cols=np.concatenate([cols[rp[r]:rp[r + 1]] for r in rows]),
                              sample_vals=np.concatenate([vals[rp[r]:rp[r + 1]] for r in rows]),
                              spmm_y_sample=Y[rows])
                meta["spmm_y_sha256"] = hashlib.sha256(np.ascontiguousarray(Y.T).tobytes()).hexdigest()
            np.savez_compressed(os.path.join(OUT, name + ".npz"), **arrays)
            with open(os.path.join(OUT, name + ".json"), "w") as f:
                json.dump(meta, f, indent=1)
            print(name, n, info["nnz_upper"], E[:3])
        finally:
            shutil.rmtree(d)


def hh_only():
    """Adds the reference's .HH row product (SparseMatrix::TimesMatrix, type 1) on the
    deterministic 3-vector input: hh_y in the npz for the full cases, its SHA-256 and the
    sampled rows otherwise."""
    for (name, shell, nv, zv, twojz, par, seed, full, bloc) in CASES:
        d = tempfile.mkdtemp()
        try:
            synth.write_system(d, shell, name, nv, zv, twojz, par, seed)
            run([os.path.join(REF, "XsysmbsOMP"), name + ".xml"], d)
            run([os.path.join(REF, "XHH+JJ"), name + ".xml"], d)
            run([os.path.join(REF, "ref_harness"), "hhtimes", name + ".HH", "3", "hh.bin"], d)
            raw = open(os.path.join(d, "hh.bin"), "rb").read()
            n, m = struct.unpack_from("<QQ", raw, 0)
            Y = np.frombuffer(raw, np.float64, n * m, 16).reshape(m, n).T
            npz = os.path.join(OUT, name + ".npz")
            arrays = dict(np.load(npz))
            jp = os.path.join(OUT, name + ".json")
            meta = json.load(open(jp))
            meta["hh_y_sha256"] = hashlib.sha256(np.ascontiguousarray(Y.T).tobytes()).hexdigest()
            if full:
                arrays["hh_y"] = Y
            else:
                arrays["hh_y_sample"] = Y[arrays["sample_rows"]]
            np.savez_compressed(npz, **arrays)
            with open(jp, "w") as f:
                json.dump(meta, f, indent=1)
            print(name, "hh product", meta["hh_y_sha256"][:16])
        finally:
            shutil.rmtree(d)


# (case, bloc, states) of the reference Davidson runs whose counters are kept
TRAJ = [("sd2p2n_s11", 5, 5), ("sd2p2n_s11", 10, 10), ("sd2p2n_s11", 16, 16), ("sd3p1n_m1_s7", 6, 6),
        ("sd3p1n_m1_s7", 3, 3), ("fp2p2n_s5", 10, 10), ("fp2p2n_s5", 4, 4), ("fp3p2n_m1_s3", 8, 8),
        ("sd4p4n_s0", 10, 10), ("sd4p4n_s0", 5, 5)]


def traj_only():
    """dav_traj.json: the reference Davidson (DavidsonDiagonalization's settings) on the
    CSB matrix of golden cases: matrix-vector products, error flag, and every pair's
    eigenvalue, |dE| and residual."""
    cases = {c[0]: c for c in CASES}
    out = []
    for (name, bloc, nstate) in TRAJ:
        (_, shell, nv, zv, twojz, par, seed, _, _) = cases[name]
        d = tempfile.mkdtemp()
        try:
            synth.write_system(d, shell, name, nv, zv, twojz, par, seed)
            run([os.path.join(REF, "XsysmbsOMP"), name + ".xml"], d)
            r = last_json(run([os.path.join(REF, "ref_harness"), "davtraj", name + ".xml", str(bloc), "0",
                               str(nstate)], d))
            r["name"] = name
            out.append(r)
            print(name, bloc, r["nmv"], r["ierr"], max(p[2] for p in r["pairs"]))
        finally:
            shutil.rmtree(d)
    with open(os.path.join(OUT, "dav_traj.json"), "w") as f:
        json.dump(out, f, indent=1)


# Hamiltonian add-ons of <system><matrix>: name, base case, attributes
MATRIX_CASES = [("sd2p2n_s11_jjtt", "sd2p2n_s11", {"JJ": "0.5", "TT": "0.3"}),
                ("fp2p2n_s5_jj", "fp2p2n_s5", {"JJ": "-0.25"}),
                ("sd3p1n_m1_s7_tt", "sd3p1n_m1_s7", {"TT": "1.5"})]


def matrix_only():
    """<name>.xml/.int and <name>.npz for the JJ / TT cases: the reference's m-scheme
    tensors after xReadSPInt and the .HH rows XHH+JJ writes."""
    cases = {c[0]: c for c in CASES}
    for (name, base, attrs) in MATRIX_CASES:
        (_, shell, nv, zv, twojz, par, seed, _, _) = cases[base]
        d = tempfile.mkdtemp()
        try:
            xml = synth.write_system(d, shell, name, nv, zv, twojz, par, seed, matrix_attrs=attrs)
            shutil.copy(os.path.join(d, name + ".int"), os.path.join(OUT, name + ".int"))
            shutil.copy(xml, os.path.join(OUT, name + ".xml"))
            run([os.path.join(REF, "XsysmbsOMP"), name + ".xml"], d)
            run([os.path.join(REF, "ref_harness"), "dumpint", name + ".xml", "dump.bin"], d)
            (ek, ev), (vk, vv) = read_dump(os.path.join(d, "dump.bin"))
            run([os.path.join(REF, "XHH+JJ"), name + ".xml"], d)
            rp, cols, vals = oracle.read_hh(os.path.join(d, name + ".HH"))
            np.savez_compressed(os.path.join(OUT, name + ".npz"), esp_keys=ek, esp_vals=ev, vspx_keys=vk,
                                vspx_vals=vv, hh_rowptr=rp, hh_cols=cols, hh_vals=vals)
            with open(os.path.join(OUT, name + ".json"), "w") as f:
                json.dump(dict(name=name, base=base, matrix_attrs=attrs, n=int(len(rp) - 1), hh_nnz=int(rp[-1]),
                               hh_sha256=sha(os.path.join(d, name + ".HH"))), f, indent=1)
            print(name, len(rp) - 1, rp[-1], len(vv))
        finally:
            shutil.rmtree(d)


# state labelling (XSHLJT): name, base case, <matrix> attributes, <system> attributes, states
LABEL_CASES = [("sd2p2n_s11_lab", "sd2p2n_s11", {"JJ": "0.5", "TT": "0.3"}, {}, 10),
               ("sd3p1n_m1_s7_labj", "sd3p1n_m1_s7", {"TT": "1.5"}, {"J": "2"}, 6),
               ("fp2p2n_s5_labt", "fp2p2n_s5", {"JJ": "-0.25"}, {"T": "1"}, 10)]


def xml_states(path):
    """The <state> nodes of the system, attributes in file order, with their wavefunctions."""
    import xml.etree.ElementTree as ET
    out = []
    for st in ET.parse(path).getroot().find("system").findall("state"):
        out.append(dict(attrs=list(st.attrib.items()),
                        wavefunctions=[list(w.attrib.items()) for w in st.findall("wavefunction")]))
    return out


def labels_only():
    """<name>_in.xml (the system as XSHLJT reads it: after XsysmbsOMP and XHH+JJ), <name>.int,
    <name>_EE.bin (the .EE davidson_vfile wrote) and <name>_states.json (the <state> nodes
    the reference XSHLJT added from it)."""
    cases = {c[0]: c for c in CASES}
    for (name, base, mattrs, sattrs, nst) in LABEL_CASES:
        (_, shell, nv, zv, twojz, par, seed, _, _) = cases[base]
        d = tempfile.mkdtemp()
        try:
            synth.write_system(d, shell, name, nv, zv, twojz, par, seed, matrix_attrs=mattrs, system_attrs=sattrs)
            shutil.copy(os.path.join(d, name + ".int"), os.path.join(OUT, name + ".int"))
            run([os.path.join(REF, "XsysmbsOMP"), name + ".xml"], d)
            run([os.path.join(REF, "XHH+JJ"), name + ".xml"], d)
            shutil.copy(os.path.join(d, name + ".xml"), os.path.join(OUT, name + "_in.xml"))
            run([os.path.join(REF, "davidson_vfile"), name + ".HH", "-b", str(nst), "-n", str(nst)], d)
            shutil.copy(os.path.join(d, name + ".HH.EE"), os.path.join(OUT, name + "_EE.bin"))
            run([os.path.join(REF, "XSHLJT"), name + ".xml", "-n", str(nst)], d)
            states = xml_states(os.path.join(d, name + ".xml"))
            with open(os.path.join(OUT, name + "_states.json"), "w") as f:
                json.dump(dict(name=name, base=base, matrix_attrs=mattrs, system_attrs=sattrs, states_asked=nst,
                               states=states), f, indent=1)
            print(name, len(states), [dict(x["attrs"])["name"] for x in states])
        finally:
            shutil.rmtree(d)


# truncated bases (XsysmbsOMP's hw and occupation rejections): name, shell, Nv, Zv, 2Jz,
# parity, seed, <system> attributes, <system> children
TRUNC_CASES = [
    ("psd2p2n_hw2", "psd", 2, 2, 0, "+", 4, {}, '<hwrejection HWMAX="2" HWMIN="0"/>'),
    ("psd2p2n_neg_hw1", "psd", 2, 2, 0, "-", 4, {}, '<hwrejection HWMAX="1" HWMIN="0"/>'),
    ("psd2p2n_hw22", "psd", 2, 2, 0, "+", 4, {}, '<hwrejection HWMAX="2" HWMIN="2"/>'),
    ("psd3p2n_m1_hwmax2", "psd", 2, 3, 1, "-", 9, {"hwmax": "2"}, ""),
    ("sd3p3n_rej", "sd", 3, 3, 0, "+", 6, {},
     '<rejection NMIN="0" NMAX="3">\n<orbital index="1" type="pn"/>\n</rejection>\n'
     '<rejection NMIN="1" NMAX="4">\n<orbital index="2"/>\n<orbital index="3" type="p"/>\n</rejection>'),
    ("fp3p2n_m1_rej", "fp", 2, 3, 1, "-", 3, {},
     '<rejection NMIN="0" NMAX="1">\n<orbital index="3" type="pn"/>\n<orbital index="4" type="pn"/>\n</rejection>'),
]


def trunc_only():
    """<name>.xml/.int, <name>.json (dimension and SHA-256 of the .mbs XsysmbsOMP writes,
    size and SHA-256 of the .HH) and <name>.npz (the .HH rows XHH+JJ writes on that
    truncated basis: all of them up to n = 4100, 64 sampled rows above)."""
    for (name, shell, nv, zv, twojz, par, seed, sattrs, kids) in TRUNC_CASES:
        d = tempfile.mkdtemp()
        try:
            xml = synth.write_system(d, shell, name, nv, zv, twojz, par, seed, system_attrs=sattrs,
                                     system_children=kids)
            shutil.copy(os.path.join(d, name + ".int"), os.path.join(OUT, name + ".int"))
            shutil.copy(xml, os.path.join(OUT, name + ".xml"))
            run([os.path.join(REF, "XsysmbsOMP"), name + ".xml"], d)
            run([os.path.join(REF, "XHH+JJ"), name + ".xml"], d)
            rp, cols, vals = oracle.read_hh(os.path.join(d, name + ".HH"))
            mbs = oracle.read_mbs(os.path.join(d, name + ".mbs"))
            if len(mbs) <= 4100:
                np.savez_compressed(os.path.join(OUT, name + ".npz"), hh_rowptr=rp, hh_cols=cols, hh_vals=vals)
            else:  # 64 sampled rows
                rows = np.sort(np.random.default_rng(2).choice(len(mbs), 64, replace=False))
                np.savez_compressed(os.path.join(OUT, name + ".npz"), sample_rows=rows,
                                    sample_rowptr=np.concatenate([[0], np.cumsum(rp[rows + 1] - rp[rows])]),
                                    sample_cols=np.concatenate([cols[rp[r]:rp[r + 1]] for r in rows]),
                                    sample_vals=np.concatenate([vals[rp[r]:rp[r + 1]] for r in rows]))
            with open(os.path.join(OUT, name + ".json"), "w") as f:
                json.dump(dict(name=name, shell=shell, nv=nv, zv=zv, twojz=twojz, parity=par, seed=seed,
                               system_attrs=sattrs, n=int(len(mbs)), mbs_sha256=sha(os.path.join(d, name + ".mbs")),
                               hh_nnz=int(rp[-1]), hh_sha256=sha(os.path.join(d, name + ".HH"))), f, indent=1)
            print(name, len(mbs), rp[-1])
        finally:
            shutil.rmtree(d)


if __name__ == "__main__":
    modes = {"csb": csb_only, "hh": hh_only, "traj": traj_only, "matrix": matrix_only, "labels": labels_only,
             "trunc": trunc_only}
    if sys.argv[1:2] and sys.argv[1] in modes:
        modes[sys.argv[1]]()
    else:
        main()
