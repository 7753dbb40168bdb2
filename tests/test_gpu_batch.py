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


def test_batch_equals_alone(gpu):
    xmls = random_sd_systems(12, 5)
    systems = [gpu.System.from_xml(x) for x in xmls]
    systems = [s for s in systems if s.dim > 0]
    nume = [min(5, s.dim) for s in systems]
    b = gpu.Batch(systems)
    evb, rsb, st = b.solve(nume, critr=1e-8)
    assert st["converged"]
    for s, k, e, r in zip(systems, nume, evb, rsb):
        op = gpu.Operator.build(s)
        ea, _, ra, _, _ = gpu.davidson(op, k, want_vectors=False, critr=1e-8, lim=k + max(20, k), maxiter=4000)
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
                                  lim=k + max(20, k), maxiter=4000)[0])
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
