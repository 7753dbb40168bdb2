"""Benchmark of the MI355X configuration-interaction hot path.

Workload (one step): one stored-H product Y = H X of the fp-shell 4p4n M=0 system
(f7/2 p3/2 f5/2 p1/2, 20 m-orbitals per species, dimension 1,963,461, seeded synthetic
isospin interaction) with k = 16 FP32 vectors, H built on the GPU beforehand and X
resident in HBM.  value = algorithmic H.X bandwidth in GB/s:
    B(k) = nnz_upper * (4 + 2 + 2) + n * k * (4 + 4)      (FP32 value, two u16 CSB indices,
                                                         X read once, Y written once)
Reported beside it:
  * hv_virtual: the same product with the proton-neutron product-form operator (no H
    stored), FP32 and FP64, with its B(k)-equivalent rate and the difference to the
    stored product;
  * roofline of the stored kernel (HIP events on its stream; HBM traffic from rocprofv3
    counters when --traffic-json names their summary);
  * solve: 10 lowest eigenpairs, block 16, all-FP64 and mixed precision, on both
    operators at basis limits 60 and 100, with the per-eigenvalue difference (MeV), the
    fastest-FP64 over fastest-mixed ratio and solve rates;
  * solve_sd6p6n: sd 6p6n (n = 93,710), block 8, 10 pairs, stored, FP64 against mixed at
    basis limits 30 and 90;
  * solve_ref16: the reference's reading of block 16 / 10 states (16 pairs targeted, the
    lowest 10 kept), product form, FP64 against mixed at basis limits 100 and 128;
  * batch: 256 independent sd-shell solves of uneven size (Nv, Zv ~ U{1..6}, 2Jz from
    the allowed range, own seed each), mixed precision, 5 pairs each;
  * serve: a Poisson request stream (sd 70 % / fp 30 %, fp up to 4+4, k ~ U{1..10},
    tolerance log-uniform in [1e-6, 1e-4]) through the admission-controlled server,
    p50/p99 latency, solves/s and each result's difference to an FP64 solve alone;
  * cpu_baseline: the reference's own CSB product (SparseMatrixCSB::Read + TimesMatrix,
    OpenMP, FP64) on the whole matrix on this host's cores, run from oracle/_ref when that
    build is present; cpu_baseline_port: the CSB product restated in C with OpenMP on a
    row sample of the same matrix.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-solve] [--no-batch]
                       [--no-serve] [--no-cpu] [--layout auto|rows|tiled] [--calibrate-copy]
                       [--profile-solve fp64|mixed]
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "H·V GB/s (% HBM peak) + solves/s, p99 latency; mixed vs FP64 at ≤1 keV"
HBM_PEAK_GBS = 8000.0


def timed_applies(op, X, Y, prec, stream, warmup, steps):
    """Wall time per apply and per-apply event times, all on `stream`."""
    import numpy as np
    import torch
    for _ in range(warmup):
        op.apply_device(X, Y, prec, stream)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for (a, b) in evs:
        a.record(stream)
        op.apply_device(X, Y, prec, stream)
        b.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    return wall, np.array([a.elapsed_time(b) for (a, b) in evs])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--shell", default="fp")
    ap.add_argument("--nv", type=int, default=4)
    ap.add_argument("--zv", type=int, default=4)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-virtual", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=0, help="CPU baseline row sample (0: n/8)")
    ap.add_argument("--layout", choices=["auto", "rows", "tiled"], default="auto",
                    help="stored-H kernel: auto (tiled from n = 262144), rows (one wave per row) or tiled")
    ap.add_argument("--traffic-json", default=None,
                    help="HBM bytes per launch of the stored kernel from rocprofv3 counters "
                         "(default: the committed summary for this kernel and workload, if any)")
    ap.add_argument("--profile-solve", choices=["fp64", "mixed"], default=None,
                    help="only one 10-pair product-form solve (block 16, basis limit 100) in this precision, "
                         "after one warm-up solve; for kernel profiles")
    ap.add_argument("--profile-batch", choices=["split", "one"], default=None,
                    help="only the 256-solve batch, after one warm-up call: default split, or every solve "
                         "in the one persistent launch; for kernel profiles")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 4 GiB device copy three times (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (the H.V path is measured on one MI355X); "
              "multi-GPU runs are not part of it", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import cosmo_amd as ca
    from cosmo_amd import synth

    torch.cuda.init()
    if args.calibrate_copy:
        x = torch.ones(1 << 30, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": x.numel() * 4, "launches": 3}))
        return

    if args.profile_batch:
        one = args.profile_batch == "one"
        batch_section(ca, synth, one_launch=one, warm=True, also_one=False)
        out = batch_section(ca, synth, one_launch=one, warm=False, also_one=False)
        print(json.dumps({"batch": args.profile_batch, "seconds": out["solve_s"], "solves": out["solves"],
                          "max_resid": out["max_resid"], "converged": out["converged"]}))
        return
    tmp = tempfile.mkdtemp(prefix="cosmo_bench_")
    name = "%s%dp%dn" % (args.shell, args.zv, args.nv)
    xml = synth.write_system(tmp, args.shell, name, args.nv, args.zv, 0, "+", args.seed)
    if args.profile_solve:
        sysm = ca.System.from_xml(xml)
        fop = ca.Operator.build(sysm, ca.OP_FACTORED)
        prec = ca.PREC_FP64 if args.profile_solve == "fp64" else ca.PREC_FP32
        kw = dict(mblock=16, lim=100, want_vectors=False, critr=1e-6, precision=prec)
        ca.davidson(fop, 10, **kw)  # warm-up: tuning and captured graphs
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev, de, rs, _, st = ca.davidson(fop, 10, **kw)
        dt = time.perf_counter() - t0
        print(json.dumps({"solve": args.profile_solve, "seconds": round(dt, 3), "iterations": st["iterations"],
                          "matvecs": st["matvecs"], "matvecs_fp64": st["matvecs_fp64"],
                          "max_resid": float(max(rs)), "evals": [round(float(x), 8) for x in ev]}))
        return
    t0 = time.perf_counter()
    sysm = ca.System.from_xml(xml)
    t_setup = time.perf_counter() - t0
    n = sysm.dim
    print("[bench] system %s n=%d setup %.2fs" % (name, n, t_setup), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    op = ca.Operator.build(sysm)
    if args.layout != "auto":
        op.set_layout({"rows": 1, "tiled": 2}[args.layout])
    elif sysm.dim >= 262144:
        # the benchmark applies this operator many times: the tiled layout from the first
        # product on (automatic choice: after the row kernel's 256th product)
        op.set_layout(2)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    nnz_u, nnz_f = op.nnz()
    print("[bench] H built: nnz_upper=%d nnz_full=%d in %.2fs (%.2f GB on device)" %
          (nnz_u, nnz_f, t_build, op.bytes() / 1e9), file=sys.stderr, flush=True)

    k = args.k
    g = torch.Generator(device="cuda").manual_seed(1)
    X = torch.randn(n, k, device="cuda", dtype=torch.float32, generator=g)
    Y = torch.empty_like(X)
    stream = torch.cuda.Stream()
    # the first product builds the tiled layout (and, on the tiled layout, times its launch
    # configurations): reported on its own, outside the timed steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op.apply_device(X, Y, ca.PREC_FP32, stream)
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    ms_per_step, kms = timed_applies(op, X, Y, ca.PREC_FP32, stream, args.warmup, args.steps)
    bytes_alg = nnz_u * (4 + 2 + 2) + n * k * (4 + 4)
    flops = 2.0 * k * (2 * nnz_u - n)
    gbs = bytes_alg / (ms_per_step * 1e-3) / 1e9
    kavg = float(kms.mean())
    achieved = bytes_alg / (kavg * 1e-3) / 1e9
    print("[bench] stored H.V k=%d: %.3f ms/step wall, kernel avg %.3f ms -> %.1f GB/s algorithmic" %
          (k, ms_per_step, kavg, achieved), file=sys.stderr, flush=True)

    # correctness spot check of the timed product: rows of H (columns, H symmetric) times X
    rows = np.random.default_rng(0).choice(n, 8, replace=False)
    Xh = X.double().cpu().numpy()
    Yh = Y.double().cpu().numpy()
    maxrel = 0.0
    for r in rows:
        e = np.zeros((n, 1))
        e[r, 0] = 1.0
        hr = op.apply(e, ca.PREC_FP64)[:, 0]
        ref = hr @ Xh
        maxrel = max(maxrel, float(np.max(np.abs(Yh[r] - ref)) / max(1e-30, np.max(np.abs(ref)))))

    # X and Y are dense n x k, so the row pitch is k: the tiled kernel serves pitches 4, 8, 16
    tiled = k in (4, 8, 16) and (args.layout == "tiled" or (args.layout == "auto" and n >= 262144))
    # the tiled stored product is two launches per H.V: the proton-spectator part
    # (k_pm_apply) and the panels (k_sell_apply); the events and the traffic cover both
    kname = ("k_pm_apply<float,%d> + k_sell_apply<float,%d>" % (k, k)) if tiled else "k_spmm_vec<float,4>"
    kfile = ("k_sell_apply" if tiled else "k_spmm_vec") + "_%s_k%d_traffic.json" % (name, k)
    traffic = None
    tj = args.traffic_json or os.path.join(ROOT, "profiles", kfile)
    if os.path.exists(tj) and (args.traffic_json or args.seed == 2):
        tr = json.load(open(tj))
        if tr.get("kernel") == kname:
            traffic = tr.get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC, "value": round(gbs, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "%s-shell %dp%dn M=0 stored FP32 H, H.V of k=%d vectors" % (args.shell, args.zv,
                                                                                          args.nv, k),
                   "n": n, "nnz_upper": nnz_u, "nnz_full_stored": nnz_f, "k": k, "seed": args.seed,
                   "parallelism": "single GPU"},
        "hv": {"frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 4), "gflops": round(flops / (ms_per_step * 1e-3) / 1e9, 1),
               "kernel_ms_p50": round(float(np.percentile(kms, 50)), 4),
               "kernel_ms_p99": round(float(np.percentile(kms, 99)), 4), "spot_check_max_rel_err": maxrel,
               "build_s": round(t_build, 3), "setup_s": round(t_setup, 3),
               "first_product_s": round(t_first, 3), "device_bytes": op.bytes()},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": kname, "bytes_per_launch": bytes_alg},
    }

    print("[bench] spot check of %d rows: max rel err %.2e" % (len(rows), maxrel), file=sys.stderr, flush=True)
    fop = None
    if not args.no_virtual:
        fop, out["hv_virtual"] = virtual_section(ca, sysm, op, X, Y, stream, args, bytes_alg, ms_per_step)
    if not args.no_solve:
        out["solve"] = solve_section(ca, fop if fop is not None else op, op, n)
        out["solve_sd6p6n"] = solve_sd6p6n_section(ca, synth)
        out["solve_sd4p4n"] = solve_sd4p4n_section(ca, synth)
        if fop is not None:
            out["solve_ref16"] = solve_ref16_section(ca, fop)
    del X, Y
    fop = None  # frees the product-form operator
    if not args.no_batch:
        batch_section(ca, synth, warm=True, also_one=False)  # warm-up: kernel loading, pools
        out["batch"] = batch_section(ca, synth)
    if not args.no_serve:
        out["serve"] = serve_section(ca, synth)
    if not args.no_cpu:
        port = cpu_section(ca, op, n, k, args.cpu_rows)
        try:  # a side measurement: it never stops the bench
            ref = cpu_reference_section(ca, op, n, k)
        except Exception as e:  # noqa: BLE001
            print("[bench] cpu reference failed: %r" % (e,), file=sys.stderr, flush=True)
            ref = None
        out["cpu_baseline"] = ref if ref is not None else port
        out["cpu_baseline_port"] = port
    print(json.dumps(out), flush=True)


def virtual_section(ca, sysm, op, X, Y, stream, args, bytes_alg, ms_stored):
    """Product-form operator: build, FP32 and FP64 applies of the same X, agreement."""
    import numpy as np
    import torch
    t0 = time.perf_counter()
    fop = ca.Operator.build(sysm, ca.OP_FACTORED)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    Yf = torch.empty_like(X)
    ms32, k32 = timed_applies(fop, X, Yf, ca.PREC_FP32, stream, args.warmup, args.steps)
    rel32 = float((Yf - Y).abs().max() / Y.abs().max())
    X64 = X.double()
    Y64 = torch.empty_like(X64)
    ms64, k64 = timed_applies(fop, X64, Y64, ca.PREC_FP64, stream, 2, max(5, args.steps // 4))
    Ys64 = torch.empty_like(X64)
    op.apply_device(X64, Ys64, ca.PREC_FP64, stream)
    torch.cuda.synchronize()
    rel64 = float((Y64 - Ys64).abs().max() / Ys64.abs().max())
    print("[bench] product-form H.V k=%d: fp32 %.3f ms, fp64 %.3f ms (stored fp32 %.3f ms); rel diff fp32 %.1e "
          "fp64 %.1e" % (X.shape[1], ms32, ms64, ms_stored, rel32, rel64), file=sys.stderr, flush=True)
    del X64, Y64, Ys64, Yf
    return fop, {
        "operator": "proton-neutron product form (no H stored)", "build_s": round(t_build, 3),
        "device_bytes": fop.bytes(),
        "fp32": {"ms_per_step": round(ms32, 4), "event_ms_p50": round(float(np.percentile(k32, 50)), 4),
                 "equiv_GBs": round(bytes_alg / (ms32 * 1e-3) / 1e9, 1), "speedup_vs_stored": round(ms_stored / ms32, 3),
                 "max_rel_diff_vs_stored_fp32": rel32},
        "fp64": {"ms_per_step": round(ms64, 4), "event_ms_p50": round(float(np.percentile(k64, 50)), 4),
                 "max_rel_diff_vs_stored_fp64": rel64},
        "note": "equiv_GBs is B(k) of the stored product divided by this apply's time",
    }


def solve_section(ca, fop, op, n, lims=(60, 100)):
    """10 lowest pairs, block 16, all-FP64 and mixed, on the product-form operator (the
    faster one) and on the stored one, at two basis limits (a wider basis restarts less
    and helps both precisions, FP64 more).  mixed_speedup compares the fastest FP64 run
    with the fastest mixed run; mixed_speedup_same_lim[lim] compares them at each limit."""
    import numpy as np
    res = {}
    for oname, o in (("virtual", fop), ("stored", op)):
        if o is None:
            continue
        r = {}
        raw = {}
        for lim in lims:
            for label, prec in (("fp64", ca.PREC_FP64), ("mixed", ca.PREC_FP32)):
                t0 = time.perf_counter()
                ev, de, rs, _, st = ca.davidson(o, 10, mblock=16, lim=lim, want_vectors=False, critr=1e-6,
                                                precision=prec)
                dt = time.perf_counter() - t0
                raw[(label, lim)] = np.array(ev)
                r["%s_lim%d" % (label, lim)] = dict(
                    seconds=round(dt, 3), evals=[round(float(x), 8) for x in ev], max_resid=float(np.max(rs)),
                    iterations=st["iterations"], matvecs=st["matvecs"], matvecs_fp64=st["matvecs_fp64"],
                    ms_apply=round(st["ms_apply"], 1), ms_ortho=round(st["ms_ortho"], 1),
                    ms_rr=round(st["ms_rr"], 1), converged=st["converged"])
                print("[bench] solve %s %s lim %d: %.2fs iters %d resid %.2e" % (oname, label, lim, dt,
                                                                              st["iterations"], np.max(rs)),
                      file=sys.stderr, flush=True)
        best = {lab: min(lims, key=lambda L: r["%s_lim%d" % (lab, L)]["seconds"]) for lab in ("fp64", "mixed")}
        t64, t32 = (r["%s_lim%d" % (lab, best[lab])]["seconds"] for lab in ("fp64", "mixed"))
        # every mixed eigenvalue against the fastest FP64 run (and against FP64 at its own limit)
        d = np.abs(raw[("mixed", best["mixed"])] - raw[("fp64", best["fp64"])])
        d_all = max(float(np.max(np.abs(raw[("mixed", L)] - raw[("fp64", L)]))) for L in lims)
        r["best_lim"] = best
        r["abs_dE_MeV"] = [float(x) for x in d]
        r["max_abs_dE_MeV"] = float(max(d.max(), d_all))
        r["within_1keV"] = bool(r["max_abs_dE_MeV"] <= 1e-3)
        r["mixed_speedup"] = round(t64 / t32, 3)
        r["mixed_speedup_same_lim"] = {str(L): round(r["fp64_lim%d" % L]["seconds"] / r["mixed_lim%d" % L]["seconds"], 3)
                                       for L in lims}
        r["solves_per_s_mixed"] = round(1.0 / t32, 4)
        r["eigenpairs_per_s_mixed"] = round(10.0 / t32, 3)
        res[oname] = r
    return res


def solve_ref16_section(ca, fop, lims=(100, 128)):
    """The reference's reading of "block 16, 10 states": DavidsonDiagonalization with
    bloc = 16 targets 16 pairs (numemax = mblock = bloc, include/kdav/davidson.cxx:167-223)
    and saves the lowest nstates = 10.  Product-form operator, FP64 against mixed at two
    basis limits; fastest against fastest, and the error of the 10 saved eigenvalues."""
    import numpy as np
    r, raw = {"targeted_pairs": 16, "saved_pairs": 10, "block": 16}, {}
    for lim in lims:
        for label, prec in (("fp64", ca.PREC_FP64), ("mixed", ca.PREC_FP32)):
            t0 = time.perf_counter()
            ev, de, rs, _, st = ca.davidson(fop, 16, mblock=16, lim=lim, want_vectors=False, critr=1e-6,
                                            precision=prec)
            dt = time.perf_counter() - t0
            raw[(label, lim)] = np.array(ev[:10])
            r["%s_lim%d" % (label, lim)] = dict(seconds=round(dt, 3), max_resid=float(np.max(rs)),
                                                iterations=st["iterations"], matvecs=st["matvecs"],
                                                matvecs_fp64=st["matvecs_fp64"], converged=st["converged"])
            print("[bench] solve 16-targeted %s lim %d: %.2fs iters %d resid %.2e" % (label, lim, dt,
                                                                                  st["iterations"], np.max(rs)),
                  file=sys.stderr, flush=True)
    best = {lab: min(lims, key=lambda L: r["%s_lim%d" % (lab, L)]["seconds"]) for lab in ("fp64", "mixed")}
    d = np.abs(raw[("mixed", best["mixed"])] - raw[("fp64", best["fp64"])])
    r["best_lim"] = best
    r["abs_dE_MeV"] = [float(x) for x in d]
    r["max_abs_dE_MeV"] = float(d.max())
    r["within_1keV"] = bool(d.max() <= 1e-3)
    r["mixed_speedup"] = round(r["fp64_lim%d" % best["fp64"]]["seconds"] / r["mixed_lim%d" % best["mixed"]]["seconds"], 3)
    return r


def solve_sd6p6n_section(ca, synth, lims=(30, 90)):
    """sd shell 6p6n M=0 (n = 93,710, seed 1): stored FP64 Davidson against FP32 H.V with
    FP64 Rayleigh-Ritz, residuals and refinement, block 8, lowest 10 pairs, at two basis
    limits (30 = the reference's nume + 20); fastest FP64 run against fastest mixed run,
    and the per-eigenvalue error of the mixed runs against the FP64 runs."""
    import numpy as np
    tdir = tempfile.mkdtemp(prefix="cosmo_sd6_")
    s = ca.System.from_xml(synth.write_system(tdir, "sd", "sd6p6n", 6, 6, 0, "+", 1))
    op = ca.Operator.build(s)
    ca.davidson(op, 10, mblock=8, lim=lims[0], want_vectors=False, critr=1e-6, precision=ca.PREC_FP32)  # warm
    r, raw = {"n": int(s.dim), "block": 8, "pairs": 10}, {}
    for lim in lims:
        for label, prec in (("fp64", ca.PREC_FP64), ("mixed", ca.PREC_FP32)):
            t0 = time.perf_counter()
            ev, de, rs, _, st = ca.davidson(op, 10, mblock=8, lim=lim, want_vectors=False, critr=1e-6,
                                            precision=prec)
            dt = time.perf_counter() - t0
            raw[(label, lim)] = np.array(ev)
            r["%s_lim%d" % (label, lim)] = dict(seconds=round(dt, 3), max_resid=float(np.max(rs)),
                                                iterations=st["iterations"], matvecs=st["matvecs"],
                                                matvecs_fp64=st["matvecs_fp64"], converged=st["converged"])
            print("[bench] solve sd6p6n %s lim %d: %.3fs iters %d resid %.2e" % (label, lim, dt, st["iterations"],
                                                                             np.max(rs)),
                  file=sys.stderr, flush=True)
    best = {lab: min(lims, key=lambda L: r["%s_lim%d" % (lab, L)]["seconds"]) for lab in ("fp64", "mixed")}
    d = np.abs(raw[("mixed", best["mixed"])] - raw[("fp64", best["fp64"])])
    d_all = max(float(np.max(np.abs(raw[("mixed", L)] - raw[("fp64", L)]))) for L in lims)
    r["best_lim"] = best
    r["evals_fp64"] = [float(x) for x in raw[("fp64", best["fp64"])]]
    r["abs_dE_MeV"] = [float(x) for x in d]
    r["max_abs_dE_MeV"] = float(max(d.max(), d_all))
    r["within_1keV"] = bool(r["max_abs_dE_MeV"] <= 1e-3)
    r["mixed_speedup"] = round(r["fp64_lim%d" % best["fp64"]]["seconds"] / r["mixed_lim%d" % best["mixed"]]["seconds"], 3)
    r["mixed_speedup_same_lim"] = {str(L): round(r["fp64_lim%d" % L]["seconds"] / r["mixed_lim%d" % L]["seconds"], 3)
                                   for L in lims}
    op.close()
    try:  # a side measurement: it never stops the bench
        ref = cpu_reference_solve(tdir, "sd6p6n", 10)
    except Exception as e:  # noqa: BLE001
        print("[bench] cpu reference solve failed: %r" % (e,), file=sys.stderr, flush=True)
        ref = None
    if ref is not None:
        e_ref = np.array(ref.pop("evals"))
        ref.pop("resid")
        ref["max_abs_dE_vs_gpu_fp64_MeV"] = float(np.max(np.abs(e_ref - raw[("fp64", best["fp64"])])))
        # Davidson time alone on both sides (H built beforehand on both)
        ref["gpu_mixed_speedup"] = round(ref["t_davidson_s"] / r["mixed_lim%d" % best["mixed"]]["seconds"], 1)
        ref["gpu_fp64_speedup"] = round(ref["t_davidson_s"] / r["fp64_lim%d" % best["fp64"]]["seconds"], 1)
        r["cpu_reference"] = ref
    return r


def solve_sd4p4n_section(ca, synth):
    """The plain CPU reference workload: sd shell 4p4n M=0 (n = 28,503, seed 0), FP64 CSB H,
    the lowest 10 pairs with the reference driver's settings (block = states = 10, basis
    limit 30, critr 1e-8), on the GPU in FP64 and in mixed precision, beside the reference's
    own build and Davidson (PPMatrix2CSBBlk + DavidsonDiagonalization, OpenMP) on this host."""
    import numpy as np
    tdir = tempfile.mkdtemp(prefix="cosmo_sd4_")
    s = ca.System.from_xml(synth.write_system(tdir, "sd", "sd4p4n", 4, 4, 0, "+", 0))
    t0 = time.perf_counter()
    op = ca.Operator.build(s)
    t_build = time.perf_counter() - t0
    ca.davidson(op, 10, want_vectors=False, precision=ca.PREC_FP32)  # warm
    r, raw = {"n": int(s.dim), "block": 10, "pairs": 10, "lim": 30, "critr": 1e-8, "gpu_build_s": round(t_build, 3)}, {}
    for label, prec in (("fp64", ca.PREC_FP64), ("mixed", ca.PREC_FP32)):
        t0 = time.perf_counter()
        ev, de, rs, _, st = ca.davidson(op, 10, want_vectors=False, precision=prec)
        dt = time.perf_counter() - t0
        raw[label] = np.array(ev)
        r[label] = dict(seconds=round(dt, 3), max_resid=float(np.max(rs)), iterations=st["iterations"],
                        matvecs=st["matvecs"], matvecs_fp64=st["matvecs_fp64"], converged=st["converged"])
        print("[bench] solve sd4p4n %s: %.3fs iters %d resid %.2e" % (label, dt, st["iterations"], np.max(rs)),
              file=sys.stderr, flush=True)
    r["max_abs_dE_mixed_vs_fp64_MeV"] = float(np.max(np.abs(raw["mixed"] - raw["fp64"])))
    op.close()
    try:  # a side measurement: it never stops the bench
        ref = cpu_reference_solve(tdir, "sd4p4n", 10)
    except Exception as e:  # noqa: BLE001
        print("[bench] cpu reference solve failed: %r" % (e,), file=sys.stderr, flush=True)
        ref = None
    if ref is not None:
        e_ref = np.array(ref.pop("evals"))
        # pairs the reference left unconverged (residual above 1e-6) are not compared
        ok = [i for i, x in enumerate(ref.pop("resid")) if x <= 1e-6]
        ref["compared_pairs"] = len(ok)
        ref["max_abs_dE_vs_gpu_fp64_MeV"] = float(np.max(np.abs(e_ref[ok] - raw["fp64"][ok]))) if ok else None
        ref["gpu_fp64_speedup"] = round(ref["t_davidson_s"] / r["fp64"]["seconds"], 1)
        ref["gpu_mixed_speedup"] = round(ref["t_davidson_s"] / r["mixed"]["seconds"], 1)
        r["cpu_reference"] = ref
    return r


def cpu_reference_solve(d, name, bloc):
    """The reference's own path on this host for the system in d: XsysmbsOMP writes the
    basis, then oracle/_ref/ref_harness builds the CSB H with PPMatrix2CSBBlk and runs
    DavidsonDiagonalization (block = states = bloc, its defaults: critr 1e-8, lim bloc + 20)
    with OpenMP.  None when those builds are absent."""
    import json as _json
    import subprocess
    import numpy as np
    import oracle
    ref = os.path.join(ROOT, "oracle", "_ref")
    mbs, exe = os.path.join(ref, "XsysmbsOMP"), os.path.join(ref, "ref_harness")
    if not (os.access(mbs, os.X_OK) and os.access(exe, os.X_OK)):
        return None
    try:
        subprocess.run([mbs, name + ".xml"], cwd=d, capture_output=True, text=True, timeout=120, check=True)
        t0 = time.perf_counter()
        out = subprocess.run([exe, "csb", name + ".xml", "ref.EE", str(bloc), "0", str(bloc)], cwd=d,
                             capture_output=True, text=True, timeout=150, check=True).stdout
        wall = time.perf_counter() - t0
    except (subprocess.SubprocessError, OSError) as e:
        print("[bench] cpu reference solve skipped: %s" % e, file=sys.stderr, flush=True)
        return None
    info = _json.loads([ln for ln in out.splitlines() if ln.startswith("{")][-1])
    E, dE, R, V = oracle.read_ee(os.path.join(d, "ref.EE"))
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    print("[bench] cpu reference solve: build %.2fs, Davidson %.2fs" % (info["t_build"], info["t_davidson"]),
          file=sys.stderr, flush=True)
    return {"cores": cores, "block": bloc, "pairs": bloc, "critr": 1e-8, "t_build_s": round(info["t_build"], 3),
            "t_davidson_s": round(info["t_davidson"], 3), "wall_s": round(wall, 3),
            "max_resid": float(np.max(R)) if len(R) else None, "evals": [float(x) for x in E[:bloc]],
            "resid": [float(x) for x in R[:bloc]]}


def batch_section(ca, synth, one_launch=False, warm=False, also_one=True):
    """256 independent sd-shell solves of uneven size in one batch, mixed precision: the
    default split (solves up to n = 16384 in one persistent launch, the rest host-driven
    beside it), then, unless also_one is False, all 256 in the one launch."""
    import numpy as np
    t0 = time.perf_counter()
    reqs, _ = synth.random_requests(ca, 256, 17)
    systems = [r[5] for r in reqs]
    nume = [min(5, s.dim) for s in systems]
    t_setup = time.perf_counter() - t0
    print("[bench] batch: %d systems set up in %.2fs, building operators" % (len(systems), t_setup),
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b = ca.Batch(systems)
    if one_launch:
        b.set_single_launch(1 << 30)
    t_build = time.perf_counter() - t0
    print("[bench] batch: operators built in %.2fs, solving" % t_build, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ev, rs, st = b.solve(nume, critr=1e-6, precision=ca.PREC_FP32)
    dt = time.perf_counter() - t0
    b.close()
    dims = np.array([s.dim for s in systems])
    maxres = float(max(np.max(r) for r in rs))
    print("[bench] batch: 256 solves (n %d..%d) built %.2fs solved %.2fs%s" %
          (dims.min(), dims.max(), t_build, dt, " (warm-up)" if warm else ""), file=sys.stderr, flush=True)
    one = None
    if also_one:
        b1 = ca.Batch(systems)
        b1.set_single_launch(1 << 30)
        t0 = time.perf_counter()
        _, rs1, st1 = b1.solve(nume, critr=1e-6, precision=ca.PREC_FP32)
        one = {"solve_s": round(time.perf_counter() - t0, 3), "converged": st1["converged"],
               "max_resid": float(max(np.max(r) for r in rs1))}
        b1.close()
        print("[bench] batch: all 256 in one launch %.2fs" % one["solve_s"], file=sys.stderr, flush=True)
    return {"solves": len(systems), "dims_min": int(dims.min()), "dims_median": int(np.median(dims)),
            "dims_max": int(dims.max()), "eigenpairs": int(sum(nume)), "setup_s": round(t_setup, 3),
            "build_s": round(t_build, 3), "solve_s": round(dt, 3), "solves_per_s": round(len(systems) / dt, 2),
            "eigenpairs_per_s": round(sum(nume) / dt, 2), "max_resid": maxres, "converged": st["converged"],
            "precision": "mixed (FP32 H.V, FP64 refinement)",
            "one_launch": "all solves up to n = 16384 in one persistent launch, the rest host-driven beside it",
            "all_in_one_launch": one}


def serve_section(ca, synth, count=48, rate=12.0):
    """Poisson arrivals through the server; latency, throughput and FP64 agreement."""
    import numpy as np
    reqs, rng = synth.random_requests(ca, count, 23, p_sd=0.7, max_sd=6, max_fp=4)
    ks = [int(min(rng.integers(1, 11), r[5].dim)) for r in reqs]
    tols = [float(10 ** rng.uniform(-6, -4)) for _ in reqs]
    gaps = rng.exponential(1.0 / rate, size=len(reqs))
    print("[bench] serve: %d requests at %.1f/s" % (len(reqs), rate), file=sys.stderr, flush=True)
    srv = ca.Server(hbm_budget=96 << 30, max_active=8, precision=ca.PREC_FP32)
    ids = []
    t0 = time.perf_counter()
    for (r, k, tol, gap) in zip(reqs, ks, tols, gaps):
        time.sleep(float(gap))
        ids.append(srv.submit(r[5], k, tol))
    res = srv.drain()
    wall = time.perf_counter() - t0
    srv.close()
    byid = {x["id"]: x for x in res}
    lat = np.array([byid[i]["t_done"] - byid[i]["t_submit"] for i in ids])
    queue = np.array([byid[i]["t_admit"] - byid[i]["t_submit"] for i in ids])
    ok = all(byid[i]["status"] == 0 for i in ids)
    # accuracy of each result against an all-FP64 solve of the same request alone
    dE = []
    for i, (r, k, tol) in zip(ids, zip(reqs, ks, tols)):
        s = r[5]
        if s.dim > 100000:
            continue
        o = ca.Operator.build(s)
        e64 = ca.davidson(o, k, want_vectors=False, critr=tol, lim=k + max(20, k), maxiter=4000)[0]
        dE.append(float(np.max(np.abs(byid[i]["evals"] - e64))))
        del o
        if len(dE) % 8 == 0:
            print("[bench] serve: %d results checked against FP64" % len(dE), file=sys.stderr, flush=True)
    print("[bench] serve: %d requests in %.2fs, p50 %.3fs p99 %.3fs" % (len(ids), wall, np.percentile(lat, 50),
                                                                        np.percentile(lat, 99)),
          file=sys.stderr, flush=True)
    shells = [r[0] for r in reqs]
    return {"requests": len(ids), "rate_per_s": rate, "sd_fraction": round(shells.count("sd") / len(shells), 3),
            "dims_max": int(max(r[5].dim for r in reqs)), "wall_s": round(wall, 3),
            "solves_per_s": round(len(ids) / wall, 3), "eigenpairs_per_s": round(sum(ks) / wall, 3),
            "latency_p50_s": round(float(np.percentile(lat, 50)), 4),
            "latency_p99_s": round(float(np.percentile(lat, 99)), 4),
            "queue_p50_s": round(float(np.percentile(queue, 50)), 4), "all_ok": ok,
            "checked_vs_fp64": len(dE), "max_abs_dE_vs_fp64_MeV": max(dE) if dE else None,
            "within_1keV": bool(dE and max(dE) <= 1e-3), "precision": "mixed"}


def cpu_section(ca, op, n, k, rows):
    """The reference's CPU algorithm (CSB product, FP64, OpenMP over block rows) restated in
    oracle/cosmo_oracle.c, timed on this host on the first `rows` rows of the same H."""
    import ctypes
    import numpy as np
    import oracle
    L = oracle.lib()
    R = rows if rows > 0 else max(1, n // 8)
    rp, c, v = op.upper_rows(0, R)
    br = int(np.sqrt(1.0 * n)) + 1
    rpf = np.concatenate([rp, np.full(n - R, rp[-1])]).astype(np.int64)
    h = L.orc_csb_build(n, br, oracle._p(rpf), oracle._p(c), oracle._p(v))
    X = np.random.default_rng(0).standard_normal((n, k))
    Y = np.zeros((n, k))
    L.orc_csb_apply(h, n, k, oracle._p(X), oracle._p(Y))  # warm-up
    reps = 0
    t0 = time.perf_counter()
    while True:
        L.orc_csb_apply(h, n, k, oracle._p(X), oracle._p(Y))
        reps += 1
        if time.perf_counter() - t0 > 8.0 or reps >= 20:
            break
    dt = (time.perf_counter() - t0) / reps
    L.orc_csb_free(ctypes.c_void_p(h))
    nnz = int(rp[-1])
    b = nnz * (8 + 2 + 2) + R * k * (8 + 8)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(b / dt / 1e9, 3), "unit": "GB/s", "cores": cores, "kind": "port",
            "sample": "rows 0..%d (%d upper entries) of the same H, FP64 CSB product, k=%d, %d reps, %.3f s/apply"
                      % (R, nnz, k, reps, dt)}


def cpu_reference_section(ca, op, n, k, reps=3):
    """The reference's own CSB product on the host: the whole H written as its CSB file
    (the layout of SparseMatrixCSB::Write), read back by the reference's
    SparseMatrixCSB::Read and multiplied by its TimesMatrix (OpenMP over block rows), in
    oracle/_ref/ref_harness, compiled from the CoSMo sources.  None when that binary is
    absent."""
    import json as _json
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.access(exe, os.X_OK):
        return None
    d = tempfile.mkdtemp(prefix="cosmo_csb_")
    path = os.path.join(d, "h.csb")
    try:
        t0 = time.perf_counter()
        try:
            op.write_csb(path)
        except ca.CosmoError as e:  # e.g. no room for the file: report the port instead
            print("[bench] cpu reference skipped: %s" % e, file=sys.stderr, flush=True)
            return None
        t_write = time.perf_counter() - t0
        print("[bench] cpu reference: CSB file written (%.1f GB) in %.1fs" % (os.path.getsize(path) / 1e9, t_write),
              file=sys.stderr, flush=True)
        try:
            r = subprocess.run([exe, "csbtimes", path, str(k), str(reps)], capture_output=True, text=True,
                               timeout=150)
        except subprocess.TimeoutExpired:
            print("[bench] cpu reference timed out", file=sys.stderr, flush=True)
            return None
        if r.returncode != 0:
            print("[bench] cpu reference failed: %s" % r.stderr[-300:], file=sys.stderr, flush=True)
            return None
        info = _json.loads([ln for ln in r.stdout.splitlines() if ln.startswith("{")][-1])
    finally:
        if os.path.exists(path):
            os.remove(path)
        os.rmdir(d)
    nnz, dt = int(info["nnz_upper"]), float(info["t_spmm"])
    b = nnz * (8 + 2 + 2) + n * k * (8 + 8)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    print("[bench] cpu reference: TimesMatrix k=%d %.3f s/apply (read %.1fs)" % (k, dt, info["t_read"]),
          file=sys.stderr, flush=True)
    return {"value": round(b / dt / 1e9, 3), "unit": "GB/s", "cores": cores, "kind": "reference",
            "sample": "the whole H (%d upper entries, FP64) as its CSB file, read by the reference's "
                      "SparseMatrixCSB::Read and multiplied by its TimesMatrix, k=%d, %d reps after one warm-up, "
                      "%.3f s/apply" % (nnz, k, info["reps"], dt)}


if __name__ == "__main__":
    main()
