"""Batched solves and the serving loop: each result is the same alone and in the mix."""
import tempfile
import time

import numpy as np
import pytest

import oracle
from cosmo_amd import synth

pytestmark = pytest.mark.gpu


def random_sd_systems(count, seed, max_part=4):
    rng = np.random.default_rng(seed)
    d = tempfile.mkdtemp()
    out = []
    for i in range(count):
        nv, zv = int(rng.integers(1, max_part + 1)), int(rng.integers(1, max_part + 1))
        A = nv + zv
        # 2Jz from the allowed range, parity of A; keep small |Jz| so spaces are not empty
        twojz = int(rng.choice([j for j in range(-4, 5) if (j - A) % 2 == 0]))
        xml = synth.write_system(d, "sd", "b%d" % i, nv, zv, twojz, "+", int(rng.integers(1 << 30)))
        out.append(xml)
    return out


def batch_lim(k, n):
    """Basis limit the batch and the server give a solve of k pairs and dimension n."""
    return min(max(k + max(20, k), min(k + 40, 45)), n)


def test_batch_equals_alone(gpu):
    xmls = random_sd_systems(12, 5)
    systems = [gpu.System.from_xml(x) for x in xmls]
    systems = [s for s in systems if s.dim > 0]
    nume = [min(5, s.dim) for s in systems]
    b = gpu.Batch(systems)
    b.set_single_launch(0)  # every solve on the stream path, as cosmo_davidson runs it alone
    evb, rsb, st = b.solve(nume, critr=1e-8)
    assert st["converged"]
    for s, k, e, r in zip(systems, nume, evb, rsb):
        op = gpu.Operator.build(s)
        ea, _, ra, _, _ = gpu.davidson(op, k, want_vectors=False, critr=1e-8, lim=batch_lim(k, s.dim), maxiter=4000)
        assert np.array_equal(e, ea), "batched result differs from the solve alone"
    # spot-check against dense eigensolves
    for s, k, e in zip(systems, nume, evb):
        if s.dim <= 800:
            op = gpu.Operator.build(s)
            w, _ = op.dense_eigen(vectors=False)
            assert np.max(np.abs(e - w[:k])) < 1e-7


def test_batch_mixed_precision_within_1kev(gpu):
    xmls = random_sd_systems(8, 11)
    systems = [s for s in (gpu.System.from_xml(x) for x in xmls) if s.dim > 5]
    b = gpu.Batch(systems)
    e64, _, _ = b.solve(5, critr=1e-7)
    e32, r32, st = b.solve(5, critr=1e-7, precision=gpu.PREC_FP32)
    for a, c, r in zip(e64, e32, r32):
        assert np.max(np.abs(a - c)) <= 1e-3
        assert np.max(r) <= 1e-7


def test_server_results_same_alone_and_in_mix(gpu):
    xmls = random_sd_systems(10, 21)
    systems = [s for s in (gpu.System.from_xml(x) for x in xmls) if s.dim > 0]
    alone = []
    for s in systems:
        k = min(4, s.dim)
        alone.append(gpu.davidson(gpu.Operator.build(s), k, want_vectors=False, critr=1e-6,
                                  lim=batch_lim(k, s.dim), maxiter=4000)[0])
    # budget that holds only a few solves at once: admission must queue the rest
    foot = [gpu.footprint(s) for s in systems]
    srv = gpu.Server(hbm_budget=int(3 * max(foot) + (64 << 20)), max_active=4, precision=gpu.PREC_FP64)
    ids = {}
    rng = np.random.default_rng(0)
    for s in systems:
        ids[srv.submit(s, min(4, s.dim), 1e-6)] = s
        time.sleep(float(rng.exponential(0.01)))  # Poisson arrivals
    res = srv.drain()
    srv.close()
    assert len(res) == len(systems)
    byid = {r["id"]: r for r in res}
    for (i, s), ea in zip(ids.items(), alone):
        r = byid[i]
        assert r["status"] == 0
        assert r["t_submit"] <= r["t_admit"] <= r["t_done"]
        assert np.array_equal(r["evals"], ea), "served result differs from the solve alone"
        assert np.max(r["resid"]) <= 1e-6


def _dense_lowest(gpu, s, k):
    w, _ = gpu.Operator.build(s).dense_eigen(vectors=False)
    return w[:k]


@pytest.mark.parametrize("pairs", [5, 12])
@pytest.mark.parametrize("precision", [0, 1])
def test_single_launch_matches_dense(gpu, precision, pairs):
    """Solves of the persistent launch (one workgroup each up to n = 2048, teams of two
    above) agree with a dense eigensolve and every residual meets critr, in FP64 and mixed
    precision; 12 pairs take the kernel's 16-wide H.X staging (more than 8 vectors per
    product)."""
    xmls = random_sd_systems(24, 31)
    systems = [s for s in (gpu.System.from_xml(x) for x in xmls) if 0 < s.dim <= 2500]
    assert len(systems) >= 10
    nume = [min(pairs, s.dim) for s in systems]
    b = gpu.Batch(systems)
    b.set_single_launch(1 << 15)
    ev, rs, st = b.solve(nume, critr=1e-7, precision=precision)
    assert st["converged"]
    for s, k, e, r in zip(systems, nume, ev, rs):
        w = _dense_lowest(gpu, s, k)
        assert np.max(np.abs(e - w)) < 1e-7, (s.dim, e, w)
        assert np.max(r) <= 1e-7
    b.close()


def test_single_launch_same_alone_and_in_mix(gpu):
    """A solve of the persistent launch does not depend on the other solves of the launch,
    whether one workgroup runs it or a team (n > 2048: two workgroups or more)."""
    xmls = random_sd_systems(16, 47)
    systems = [s for s in (gpu.System.from_xml(x) for x in xmls) if 0 < s.dim <= 6000]
    assert any(s.dim > 2048 for s in systems) and any(s.dim <= 2048 for s in systems)
    nume = [min(4, s.dim) for s in systems]
    b = gpu.Batch(systems)
    b.set_single_launch(1 << 15)
    ev, rs, _ = b.solve(nume, critr=1e-6, precision=gpu.PREC_FP32)
    b.close()
    picks = {0, len(systems) // 2, len(systems) - 1, int(np.argmax([s.dim for s in systems]))}
    for i in sorted(picks):
        one = gpu.Batch([systems[i]])
        one.set_single_launch(1 << 15)
        e1, r1, _ = one.solve([nume[i]], critr=1e-6, precision=gpu.PREC_FP32)
        one.close()
        assert np.array_equal(e1[0], ev[i]) and np.array_equal(r1[0], rs[i])


def test_single_launch_tiny_and_full_space(gpu):
    """Spaces smaller than the requested count (padding) and equal to it (the basis is the
    whole space from the start): eigenvalues and residuals still exact."""
    d = tempfile.mkdtemp()
    cases = [("sd", 1, 1, 6), ("sd", 1, 1, 10), ("sd", 1, 1, 8), ("sd", 1, 2, 9)]
    systems = []
    for i, (sh, nv, zv, tj) in enumerate(cases):
        s = gpu.System.from_xml(synth.write_system(d, sh, "t%d" % i, nv, zv, tj, "+", 100 + i))
        if s.dim > 0:
            systems.append(s)
    assert any(s.dim < 5 for s in systems)
    b = gpu.Batch(systems)
    b.set_single_launch(1 << 15)
    ev, rs, st = b.solve([5] * len(systems), critr=1e-8)
    assert st["converged"]
    for s, e, r in zip(systems, ev, rs):
        k = min(5, s.dim)
        w = _dense_lowest(gpu, s, k)
        assert np.max(np.abs(e[:k] - w)) < 1e-9
        assert np.max(r[:k]) <= 1e-8
        assert np.all(e[k:] == 0.0)
    b.close()


def test_server_runs_unstorable_h_on_the_product_form(gpu, tmp_path):
    """A request whose stored H does not fit the server's HBM budget, but whose product form
    (no H in HBM) does, is served on the product form: its eigenvalues equal a stored solve
    alone to 1e-9 and its residuals meet the tolerance.  A request that fits neither way is
    refused at submission, and a one-species system has no product form."""
    xml = synth.write_system(str(tmp_path), "fp", "u", 3, 3, 0, "+", 4)
    s = gpu.System.from_xml(xml)
    k = 4
    stored, fact = gpu.footprint(s), gpu.footprint(s, gpu.OP_FACTORED)
    assert 0 < 3 * fact < stored  # the budget below holds the product form and its blocks only
    lim = batch_lim(k, s.dim)
    ea = gpu.davidson(gpu.Operator.build(s), k, want_vectors=False, critr=1e-8, lim=lim, maxiter=4000)[0]
    srv = gpu.Server(hbm_budget=stored, max_active=2, precision=gpu.PREC_FP64)
    srv.submit(s, k, 1e-8)
    res = srv.drain()
    srv.close()
    assert len(res) == 1 and res[0]["status"] == 0
    assert np.max(np.abs(res[0]["evals"] - ea)) < 1e-9
    assert np.max(res[0]["resid"]) <= 1e-8
    small = gpu.Server(hbm_budget=1 << 20, max_active=1, precision=gpu.PREC_FP64)
    with pytest.raises(gpu.CosmoError):
        small.submit(s, k, 1e-8)
    small.close()
    one = gpu.System.from_xml(synth.write_system(str(tmp_path), "sd", "one", 0, 4, 0, "+", 2))
    with pytest.raises(gpu.CosmoError):
        gpu.footprint(one, gpu.OP_FACTORED)


@pytest.mark.parametrize("one_launch", [False, True])
def test_batch_256_uneven_sd_solves_vs_oracle(gpu, one_launch):
    """The batched workload at its stated size: 256 sd-shell solves, Nv and Zv uniform in
    1..6, 2Jz from the allowed range, each with its own seeded interaction (n 18..93,710),
    5 pairs each, FP32 H.V with FP64 refinement to residual 1e-6, in one batch call: with
    the default split (solves up to n = 16384 in one persistent launch, the rest
    host-driven beside it) and with all 256 in the one launch (teams of up to 46
    workgroups).  Every eigenvalue against the oracle's (dense / Lanczos on the oracle's
    H, tests/golden/batch256_oracle.json) within 1e-6 MeV, every residual <= 1e-6."""
    import json
    import os
    from conftest import GOLDEN
    ref = json.load(open(os.path.join(GOLDEN, "batch256_oracle.json")))
    reqs, _ = synth.random_requests(gpu, 256, 17)
    systems = [r[5] for r in reqs]
    assert [s.dim for s in systems] == [x["n"] for x in ref]
    nume = [min(5, s.dim) for s in systems]
    b = gpu.Batch(systems)
    if one_launch:
        b.set_single_launch(1 << 30)
    ev, rs, st = b.solve(nume, critr=1e-6, precision=gpu.PREC_FP32)
    b.close()
    assert st["converged"]
    worst = 0.0
    for s, e, r, x in zip(systems, ev, rs, ref):
        err = np.max(np.abs(e - np.array(x["evals"])))
        worst = max(worst, err)
        assert err <= 1e-6, (s.dim, e, x["evals"])
        assert np.max(r) <= 1e-6
    print("256 solves, worst |dE| vs oracle %.2e MeV" % worst)


def test_serving_poisson_sd_fp_mixed_stream(gpu):
    """The serving workload: Poisson arrivals of sd (70 %) and fp (30 %, at most 4+4)
    requests, Nv, Zv and 2Jz drawn as in the batch, k uniform in 1..10 pairs, residual
    tolerance log-uniform in [1e-6, 1e-4], served in mixed precision while others run.
    Each result equals the same request solved alone (same precision and parameters) and
    lies within 1 keV of an all-FP64 solve alone; every residual meets its tolerance; and
    for the requests of dimension <= 1500 every eigenvalue lies within its residual
    tolerance of the oracle's (numpy eigvalsh of the oracle's dense H: a Ritz value with
    residual r is within r of an eigenvalue)."""
    reqs, rng = synth.random_requests(gpu, 24, 23, p_sd=0.7, max_sd=6, max_fp=4)
    ks = [int(min(rng.integers(1, 11), r[5].dim)) for r in reqs]
    tols = [float(10 ** rng.uniform(-6, -4)) for _ in reqs]
    gaps = rng.exponential(1.0 / 40.0, size=len(reqs))  # arrivals faster than most solves finish
    assert any(r[0] == "fp" for r in reqs) and any(r[0] == "sd" for r in reqs)
    srv = gpu.Server(hbm_budget=96 << 30, max_active=8, precision=gpu.PREC_FP32)
    ids = []
    for r, k, tol, gap in zip(reqs, ks, tols, gaps):
        time.sleep(float(gap))
        ids.append(srv.submit(r[5], k, tol))
    res = srv.drain()
    srv.close()
    byid = {x["id"]: x for x in res}
    overlapped = checked = 0
    for i, r, k, tol in zip(ids, reqs, ks, tols):
        x = byid[i]
        assert x["status"] == 0
        assert np.max(x["resid"]) <= tol
        overlapped += any(y["t_admit"] < x["t_done"] and x["t_admit"] < y["t_done"] for j, y in byid.items() if j != i)
        s = r[5]
        lim = batch_lim(k, s.dim)
        op = gpu.Operator.build(s)
        e32 = gpu.davidson(op, k, want_vectors=False, critr=tol, lim=lim, maxiter=4000, precision=gpu.PREC_FP32)[0]
        e64 = gpu.davidson(op, k, want_vectors=False, critr=tol, lim=lim, maxiter=4000)[0]
        op.close()
        assert np.array_equal(x["evals"], e32), "served result differs from the request solved alone"
        assert np.max(np.abs(x["evals"] - e64)) <= 1e-3
        if s.dim <= 1500:
            o = oracle.System(r[4])
            w = np.linalg.eigvalsh(oracle.dense_from_upper(s.dim, *o.h_upper(basis=s.basis())))
            assert np.max(np.abs(x["evals"][:k] - w[:k])) <= tol, (s.dim, x["evals"][:k], w[:k])
            checked += 1
    assert overlapped >= len(ids) // 2  # requests really ran beside one another
    assert checked >= 5  # requests checked against the oracle
