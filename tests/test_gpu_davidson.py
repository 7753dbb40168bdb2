"""Block Davidson on the GPU against dense eigensolves of the oracle H and against the
reference Davidson's eigenvalues (where the reference's own residual is small)."""
import numpy as np
import pytest

import oracle
from conftest import FULL_CASES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,k", [("sd2p2n_s11", 8), ("sd3p1n_m1_s7", 6), ("fp2p2n_s5", 8)])
def test_lowest_pairs_vs_dense(gpu, golden, name, k):
    g = golden(name)
    s = gpu.System.from_xml(g.xml)
    op = gpu.Operator.build(s)
    o = oracle.System(g.xml)
    rp, c, v = o.h_upper()
    H = oracle.dense_from_upper(op.n, rp, c, v)
    w = np.linalg.eigvalsh(H)
    ev, de, rs, vec, st = gpu.davidson(op, k)
    assert st["converged"]
    assert np.max(np.abs(ev - w[:k])) < 1e-8
    assert np.max(rs) < 1e-6
    R = H @ vec - vec * ev
    assert np.max(np.linalg.norm(R, axis=0)) < 1e-6
    assert np.allclose(vec.T @ vec, np.eye(k), atol=1e-8)


@pytest.mark.parametrize("name", ["sd4p4n_s0", "fp3p2n_m1_s3"])
def test_reference_eigenvalues(gpu, golden, name):
    g = golden(name)
    op = gpu.Operator.build(gpu.System.from_xml(g.xml))
    k = g.meta["davidson_block"]
    # a wider basis than the reference default (nume+20): its last pair sits next to a
    # near-degenerate neighbour, where the reference itself stopped unconverged
    ev, de, rs, vec, st = gpu.davidson(op, k, want_vectors=False, lim=k + 40, maxiter=2000)
    assert st["converged"]
    ref = np.array(g.meta["ref_evals"])
    rres = np.array(g.meta["ref_resid"])
    ok = rres <= 1e-6
    assert ok.sum() >= k // 2
    assert np.max(np.abs(ev[ok] - ref[ok])) < 1e-7
    assert np.max(rs) < 1e-6


def test_mixed_precision_within_1kev(gpu, golden):
    g = golden("sd4p4n_s0")
    op = gpu.Operator.build(gpu.System.from_xml(g.xml))
    e64, _, r64, _, st64 = gpu.davidson(op, 10, want_vectors=False, critr=1e-6)
    e32, _, r32, _, st32 = gpu.davidson(op, 10, want_vectors=False, critr=1e-6, precision=gpu.PREC_FP32)
    assert np.max(np.abs(e32 - e64)) <= 1e-3
    assert np.max(r32) <= 1e-6
    assert st32["matvecs_fp64"] < st32["matvecs"]


def test_block_smaller_than_pairs(gpu, golden):
    g = golden("fp2p2n_s5")
    op = gpu.Operator.build(gpu.System.from_xml(g.xml))
    o = oracle.System(g.xml)
    rp, c, v = o.h_upper()
    w = np.linalg.eigvalsh(oracle.dense_from_upper(op.n, rp, c, v))
    # 4 new vectors per iteration for 10 pairs: the reference basis limit (nume + 20) and
    # maxiter (400) leave the last pairs above critr; a wider basis converges all of them
    ev, de, rs, vec, st = gpu.davidson(op, 10, mblock=4, want_vectors=False, lim=50, maxiter=2000)
    assert st["converged"] and rs.max() <= 1e-8
    assert np.max(np.abs(ev - w[:10])) < 1e-8


def test_virtual_operator_solve(gpu, golden):
    g = golden("sd2p2n_s11")
    s = gpu.System.from_xml(g.xml)
    a = gpu.Operator.build(s)
    b = gpu.Operator.build(s, gpu.OP_VIRTUAL)
    ea = gpu.davidson(a, 6, want_vectors=False)[0]
    eb = gpu.davidson(b, 6, want_vectors=False)[0]
    assert np.max(np.abs(ea - eb)) < 1e-9


def test_ee_roundtrip(gpu, golden, tmp_path):
    g = golden("sd2p2n_s11")
    op = gpu.Operator.build(gpu.System.from_xml(g.xml))
    ev, de, rs, vec, st = gpu.davidson(op, 4)
    p = str(tmp_path / "x.EE")
    gpu.write_ee(p, ev, de, rs, vec)
    E, dE, R, V = oracle.read_ee(p)
    assert np.array_equal(E, ev) and np.array_equal(R, rs)
    assert np.array_equal(np.stack(V, 1), vec)
    # warm start from the written vectors converges at once to the same values
    ev2 = gpu.davidson(op, 4, init=vec, want_vectors=False)[0]
    assert np.max(np.abs(ev2 - ev)) < 1e-9


@pytest.mark.parametrize("seed", [3, 17, 41])
def test_mixed_precision_full_space_basis(gpu, tmp_path, seed):
    """A basis that grows to the whole space (n = 25, 5 pairs, lim = n): the FP32 phase's
    vectors must be re-orthonormalised in FP64 before the refinement, or the last pairs
    stall above the residual target."""
    from cosmo_amd import synth
    xml = synth.write_system(str(tmp_path), "sd", "t", 1, 1, -2, "+", seed)
    op = gpu.Operator.build(gpu.System.from_xml(xml))
    assert op.n == 25
    kw = dict(want_vectors=False, critr=1e-6, lim=25, maxiter=4000)
    e64, _, r64, _, _ = gpu.davidson(op, 5, **kw)
    e32, _, r32, _, st = gpu.davidson(op, 5, precision=gpu.PREC_FP32, **kw)
    assert st["converged"]
    assert np.max(r32) <= 1e-6
    assert np.max(np.abs(e32 - e64)) < 1e-8


def test_sd6p6n_mixed_vs_fp64_block8(gpu, tmp_path):
    """sd shell 6p6n M=0 (n = 93,710): FP32 H.V with FP64 Rayleigh-Ritz, residuals and
    refinement against the all-FP64 solve, block 8, lowest 10 pairs.  Every eigenvalue within
    1 keV (in practice ~1e-10 MeV), every residual below 1e-6; the lowest pair also checked
    by the Rayleigh quotient of an independent FP64 product with the returned vector."""
    from cosmo_amd import synth
    s = gpu.System.from_xml(synth.write_system(str(tmp_path), "sd", "c1", 6, 6, 0, "+", 1))
    assert s.dim == 93710
    op = gpu.Operator.build(s)
    e64, _, r64, v64, _ = gpu.davidson(op, 10, mblock=8, lim=30, critr=1e-6)
    e32, _, r32, _, st = gpu.davidson(op, 10, mblock=8, lim=30, critr=1e-6, precision=gpu.PREC_FP32,
                                      want_vectors=False)
    assert st["converged"] and st["matvecs_fp64"] < st["matvecs"]
    assert np.max(np.abs(e32 - e64)) <= 1e-3
    assert np.max(r64) <= 1e-6 and np.max(r32) <= 1e-6
    x = v64[:, 0]
    hx = op.apply(x[:, None], gpu.PREC_FP64)[:, 0]
    assert abs(x @ hx / (x @ x) - e64[0]) < 1e-9


def _traj():
    import json
    import os
    from conftest import GOLDEN
    return json.load(open(os.path.join(GOLDEN, "dav_traj.json")))


# runs whose H.V vector count equals the reference's exactly; in the others the blocked
# orthogonalisation and this build's Rayleigh-Ritz solver round differently and the count
# moves by at most 4.1 % (590 / 605, 277 / 266, 1264 / 1299, 701 / 697)
TRAJ_EXACT = {0, 2, 3, 6, 7, 9}


@pytest.mark.parametrize("run", range(10))
def test_trajectory_against_the_reference_davidson(gpu, golden, run):
    """The reference Davidson's counters (run by oracle/_ref/ref_harness davtraj with
    DavidsonDiagonalization's settings) against this build's with the same parameters and
    the reference's convergence test (strict=0): the same error flag, the same number of
    H.V vector products (within 5 %, exactly in the runs listed above), the reference's
    false convergences (pairs accepted on their Ritz coefficients with a residual above
    1e-2) reproduced pair for pair, and every pair converged on both sides (residual
    <= 1e-8) at the reference's eigenvalue within 1e-8."""
    r = _traj()[run]
    g = golden(r["name"])
    op = gpu.Operator.build(gpu.System.from_xml(g.xml))
    k = r["bloc"]
    ev, de, rs, vec, st = gpu.davidson(op, k, want_vectors=False, strict=0, check_converged=False)
    ref = np.array(r["pairs"])
    print(r["name"], k, "matvecs", st["matvecs"], "ref", r["nmv"], "iters", st["iterations"], "resid",
          " ".join("%.1e" % x for x in rs), "ref", " ".join("%.1e" % x for x in ref[:, 2]))
    assert st["ierr"] == r["ierr"]
    assert abs(st["matvecs"] - r["nmv"]) <= 0.05 * r["nmv"]
    if run in TRAJ_EXACT:
        assert st["matvecs"] == r["nmv"]
    assert np.array_equal(rs > 1e-2, ref[:, 2] > 1e-2)
    both = (rs <= 1e-8) & (ref[:, 2] <= 1e-8)
    assert both.sum() >= (ref[:, 2] <= 1e-8).sum() - 1
    assert np.max(np.abs(ev[both] - ref[both, 0])) < 1e-8


def test_reference_defaults_unconverged_pair_is_reported(gpu, golden):
    """With the reference driver's defaults at block 5 (nume 5, lim 25, maxiter 200) the
    fifth pair of the sd 2p2n system sits 16.5 keV below its neighbour and is still at
    ~5e-5 after maxiter in the reference as here: the solve raises NotConverged (COSMO_ECONV)
    with the results attached; at the reference's block 10 every pair converges."""
    g = golden("sd2p2n_s11")
    op = gpu.Operator.build(gpu.System.from_xml(g.xml))
    with pytest.raises(gpu.NotConverged) as e:
        gpu.davidson(op, 5, want_vectors=False)
    ev, de, rs, vec, st = e.value.results
    assert not st["converged"] and st["iterations"] == 200 and rs[4] > 1e-8 and rs[:4].max() <= 1e-7
    ev, de, rs, vec, st = gpu.davidson(op, 10, want_vectors=False)
    assert st["converged"] and rs.max() <= 1e-8
    w = np.array(g.meta["ref_evals"])
    assert np.max(np.abs(ev - w)) < 1e-8


def test_strict_exit_reports_pairs_above_critr(gpu, golden):
    """Every exit of the strict solve (here the reopen rounds end on maxiter) either returns
    all residuals below critr or raises NotConverged; an unreachable critr always raises."""
    g = golden("sd3p1n_m1_s7")
    op = gpu.Operator.build(gpu.System.from_xml(g.xml))
    with pytest.raises(gpu.NotConverged) as e:
        gpu.davidson(op, 4, want_vectors=False, critr=1e-17, maxiter=60)
    ev, de, rs, vec, st = e.value.results
    assert rs.max() > 1e-17
