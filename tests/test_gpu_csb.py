"""CSB files (SparseMatrixCSB::Write / Read layout): the stored operator written as CSB
against the reference writer's file, and CSB files read back into operators."""
import hashlib
import os
import struct

import numpy as np
import pytest

import oracle
from conftest import FULL_CASES, GOLDEN_CASES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_write_csb_matches_reference_file(gpu, golden, name, tmp_path):
    """cosmo_op_write_csb of the default (1e-14-filtered) stored operator has the size,
    header, block flags, counts and indices of the file the reference's
    SparseMatrixCSB::Write made (SHA-256 of everything but the values); the values are
    the operator's own upper rows, within 1e-12 of the oracle's H."""
    g = golden(name)
    print("write_csb %s" % name, flush=True)
    op = gpu.Operator.build(gpu.System.from_xml(g.xml))
    p = str(tmp_path / "h.csb")
    op.write_csb(p)
    assert os.path.getsize(p) == g.meta["csb_bytes"]
    hdr, blocks, isha = oracle.read_csb(p)
    assert isha == g.meta["csb_index_sha256"]
    assert hdr["symmetric"] and hdr["n"] == g.meta["n"]
    assert hdr["br"] == g.meta["csb_block_rows"] and hdr["nbr"] == g.meta["csb_num_block_rows"]
    rp, c, v = oracle.csb_to_upper(hdr, blocks)
    urp, uc, uv = op.upper_rows()
    assert np.array_equal(rp, urp) and np.array_equal(c, uc) and np.array_equal(v, uv)
    orp, oc, ov = oracle.System(g.xml).h_upper(drop_small=True)
    assert np.array_equal(orp, rp) and np.array_equal(oc, c)
    assert np.max(np.abs(v - ov)) <= 1e-12 * np.abs(ov).max()


@pytest.mark.parametrize("name", FULL_CASES)
def test_read_reference_csb_file(gpu, golden, name, tmp_path):
    """The reference's CSB file (rebuilt byte for byte from its own .HH rows and checked
    against the SHA-256 of SparseMatrixCSB::Write's output) read into an operator: its
    FP64 product equals the oracle's CSB product of those rows to 1e-12."""
    g = golden(name)
    n = g.meta["n"]
    rp, c, v = g.arrays["hh_rowptr"], g.arrays["hh_cols"], g.arrays["hh_vals"]
    keep = np.abs(v) >= 1e-14
    rows = np.repeat(np.arange(n), np.diff(rp))
    frp = np.zeros(n + 1, np.int64)
    np.add.at(frp, rows[keep] + 1, 1)
    frp = np.cumsum(frp)
    b = oracle.csb_bytes(n, frp, c[keep], v[keep])
    assert hashlib.sha256(b).hexdigest() == g.meta["csb_sha256"]
    p = str(tmp_path / "ref.csb")
    open(p, "wb").write(b)
    op = gpu.Operator.read_csb(p)
    assert op.n == n and op.nnz()[0] == g.meta["nnz_upper_csb"]
    X = np.random.default_rng(3).standard_normal((n, 5))
    ref = oracle.csb_times(n, frp, c[keep], v[keep], X)
    y = op.apply(X, gpu.PREC_FP64)
    assert np.max(np.abs(y - ref)) <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("name,block", [("sd2p2n_s11", 0), ("sd2p2n_s11", 7), ("sd2p2n_s11", 1000),
                                        ("fp2p2n_s5", 0), ("fp2p2n_s5", 100), ("sd4p4n_s0", 0),
                                        ("sd4p4n_s0", 1000)])
def test_csb_round_trip_same_bits(gpu, golden, name, block, tmp_path):
    """Own file read back (default and explicit block sizes, one larger than n for the
    small systems): same upper rows, and an FP64 product bit-identical to the original
    operator's."""
    g = golden(name)
    a = gpu.Operator.build(gpu.System.from_xml(g.xml))
    p = str(tmp_path / "h.csb")
    a.write_csb(p, block)
    b = gpu.Operator.read_csb(p)
    ra, rb = a.upper_rows(), b.upper_rows()
    assert all(np.array_equal(x, y) for x, y in zip(ra, rb))
    X = np.random.default_rng(11).standard_normal((a.n, 4))
    assert np.array_equal(a.apply(X, gpu.PREC_FP64), b.apply(X, gpu.PREC_FP64))


def test_read_csb_rejects_bad_files(gpu, golden, tmp_path):
    """Truncated, non-symmetric, lower-triangle and trailing-byte files are errors."""
    g = golden("sd2p2n_s11")
    a = gpu.Operator.build(gpu.System.from_xml(g.xml))
    p = str(tmp_path / "h.csb")
    a.write_csb(p)
    good = open(p, "rb").read()
    cases = {
        "trunc": good[:-3],
        "trail": good + b"\0",
        "nonsym": b"\0" + good[1:],
    }
    # one entry of block (1, 0): a lower-triangle block
    n, br = struct.unpack_from("<IH", good, 1)
    nb = (n + br - 1) // br
    hdr = struct.pack("<?IHHHH", True, n, br, br, nb, nb)
    blocks = []
    for i in range(nb):
        for j in range(nb):
            if (i, j) == (1, 0):
                blocks.append(struct.pack("<?I", False, 1) + struct.pack("<HHd", 0, 0, 1.0))
            else:
                blocks.append(struct.pack("<?I", i == j, 0))
    cases["lower"] = hdr + b"".join(blocks)
    for k, data in cases.items():
        q = str(tmp_path / (k + ".csb"))
        open(q, "wb").write(data)
        with pytest.raises(gpu.CosmoError):
            gpu.Operator.read_csb(q)
