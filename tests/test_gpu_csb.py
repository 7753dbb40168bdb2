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


def _reference_csb(g, path):
    """The reference's CSB file rebuilt byte for byte from its own .HH rows (the SHA-256 of
    SparseMatrixCSB::Write's output is checked); returns the filtered upper rows."""
    n = g.meta["n"]
    rp, c, v = g.arrays["hh_rowptr"], g.arrays["hh_cols"], g.arrays["hh_vals"]
    keep = np.abs(v) >= 1e-14
    rows = np.repeat(np.arange(n), np.diff(rp))
    frp = np.zeros(n + 1, np.int64)
    np.add.at(frp, rows[keep] + 1, 1)
    frp = np.cumsum(frp)
    b = oracle.csb_bytes(n, frp, c[keep], v[keep])
    assert hashlib.sha256(b).hexdigest() == g.meta["csb_sha256"]
    open(path, "wb").write(b)
    return frp, c[keep], v[keep]


@pytest.mark.parametrize("name", FULL_CASES)
def test_create_csb_from_blocks(gpu, golden, name, tmp_path):
    """The reference file's blocks handed over in memory, as a SparseMatrixCSB holds them:
    the same operator as reading the file, bit for bit; shuffling the entries inside each
    block changes nothing; a repeated entry is an error."""
    g = golden(name)
    n = g.meta["n"]
    p = str(tmp_path / "ref.csb")
    _reference_csb(g, p)
    hdr, blocks, _ = oracle.read_csb(p)
    a = gpu.Operator.read_csb(p)
    blk = {(i, j): (r, c, v) for (i, j, _, r, c, v) in blocks if len(v)}
    args = (n, hdr["br"], hdr["bc"], hdr["nbr"], hdr["nbc"])
    b = gpu.Operator.from_csb(*args, blk)
    X = np.random.default_rng(2).standard_normal((n, 3))
    ya = a.apply(X, gpu.PREC_FP64)
    assert np.array_equal(b.apply(X, gpu.PREC_FP64), ya)
    rng = np.random.default_rng(9)
    shuf = {}
    for key, (r, c, v) in blk.items():
        o = rng.permutation(len(v))
        shuf[key] = (r[o], c[o], v[o])
    b2 = gpu.Operator.from_csb(*args, shuf)
    assert np.array_equal(b2.apply(X, gpu.PREC_FP64), ya)
    key = next(iter(blk))
    r, c, v = blk[key]
    dup = dict(blk)
    dup[key] = (np.concatenate([r, r[:1]]), np.concatenate([c, c[:1]]), np.concatenate([v, v[:1]]))
    with pytest.raises(gpu.CosmoError):
        gpu.Operator.from_csb(*args, dup)


@pytest.mark.parametrize("name", FULL_CASES)
def test_read_reference_csb_file(gpu, golden, name, tmp_path):
    """The reference's CSB file (rebuilt byte for byte from its own .HH rows and checked
    against the SHA-256 of SparseMatrixCSB::Write's output) read into an operator: its
    FP64 product equals the oracle's CSB product of those rows to 1e-12."""
    g = golden(name)
    n = g.meta["n"]
    p = str(tmp_path / "ref.csb")
    frp, fc, fv = _reference_csb(g, p)
    op = gpu.Operator.read_csb(p)
    assert op.n == n and op.nnz()[0] == g.meta["nnz_upper_csb"]
    X = np.random.default_rng(3).standard_normal((n, 5))
    ref = oracle.csb_times(n, frp, fc, fv, X)
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
