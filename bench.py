#!/usr/bin/env python3
"""Benchmark: one whole bcc elimination on one MI355X.

Workload: the 16384 x 65536 case-control problem (A and Z 16384 x 65536, planted
208-row bicluster in A, T/S with 3 covariate columns, gamma 0.02, B_MLT 32),
generated with a fixed seed.  A step is one elimination run to the end: the
initial loop-subscores plus every elimination pass, with the inputs already
resident in device memory (loading the .b16 files is not timed; between steps
the handle is reset to its loaded state).  The metric is the reference's
bit-operation accounting (the blop formulas of AnZt_vv.c:99, At_T_YnWt_S_Zn_ww.c:110,
An_ZtSn_ww.c:66, An_ajdk_v.c:328, xcalc.c:243) summed over the run, divided by the
elimination wall time ("elim-run s").

Output: one JSON line (see --help).  --gpus above 1 is refused: the path runs on one GPU.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_INT8_TOPS = 5000.0      # MI355X dense int8 MFMA, TOP/s
PEAK_HBM_GBS = 8000.0


def make_data(root, name, kw):
    from lakcluster_c_amd import gen
    marker = os.path.join(root, f"{name}.ready")
    in_path = os.path.join(root, f"dir_{name}_in", f"{name}.in")
    if os.path.exists(marker) and os.path.exists(in_path):
        return in_path
    t0 = time.time()
    ds = gen.make_dataset(root, name, **kw)
    with open(marker, "w") as f:
        f.write(f"{time.time() - t0:.1f}\n")
    return ds.in_path


def cpu_baseline(root, seconds_cap=30.0):
    """The CPU restatement (OpenMP) on a bounded sample of the same problem family."""
    from oracle import lko
    kw = dict(nA=1024, ncols=8192, nZ=1024, T=3, pmin=0.05, pmax=0.5, planted=(0.55,), gamma=0.02, B=32, seed=2)
    in_path = make_data(root, "cpu_sample_1024x8192", kw)
    text = open(in_path).read()
    max_iter = 40
    r = lko.run_text(text, max_iter=max_iter)
    el = r.times[1] + r.times[2]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": r.ops / el / 1e12 if el > 0 else None,
        "unit": "Tbit-op/s",
        "cores": threads,
        "kind": "port",
        "sample": f"case-control A,Z 1024x8192, T=3, gamma 0.02, B 32: initial pass + {r.n_iter} elimination "
                  f"steps, {r.ops:.3e} bit-ops in {el:.1f} s",
    }


def _traffic(workload):
    """HBM bytes per launch of the ww kernel from the committed FETCH_SIZE/WRITE_SIZE
    counter runs (copy-calibrated), or None for a workload without such a profile."""
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "c2_ww_traffic.json")
    try:
        with open(p) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload:
        return None
    return t["fetch_kib_per_launch_raw"] * 1024 * t["fetch_scale"] + t["write_kib_per_launch_raw"] * 1024 * t["write_scale"]


def tune_warmup(P, pool, seed, n_warm=32):
    """Untimed warm-up of the many-job legs: n_warm null jobs drawn with another seed from
    the same size distribution run one at a time with the autotuner measuring every new
    shape class; the store is then frozen (mode 2), so the timed plain and batched runs make
    the same choices (classes the warm-up did not see take the cost model)."""
    from lakcluster_c_amd import gen, lkc
    before = lkc.tune_get()["mode"]
    warm = gen.null_jobs(pool, n_warm, seed=seed + 100, rows=(2048, 16384), cols=(16384, 65536))
    lkc.tune_set(1)
    t = time.time()
    P.batch([(j.rowsA, j.rowsZ, j.colmask, 0.05, 32) for j in warm], max_concurrent=1)
    tw = time.time() - t
    st = lkc.tune_get()
    lkc.tune_set(2)
    print(f"[bench] tuning warm-up: {n_warm} other null jobs in {tw:.2f} s, {st['classes']} shape classes",
          file=sys.stderr, flush=True)
    return {"warmup_jobs": n_warm, "warmup_s": tw, "classes": st["classes"], "mode_before": before,
            "saved_ms_at_measurement": st["saved_ms"]}


def batched_legs(n_null=256, k_null=8, n_serve=24, k_serve=8, seed=7, c2_tbitops=None):
    """Many eliminations on one GPU, from a device-resident pool of the 16384 x 65536 case-control problem's patients
    (16384 cases + 16384 controls x 65536 SNPs, T/S 3 columns):
    * permutation nulls: row subsamples log-uniform in [2048, 16384] with re-shuffled
      case/control labels, column subsets log-uniform in [16384, 65536], gamma 0.05,
      B 32; the same jobs run one after another (the plain GPU path) and k at a time;
    * serving: a Poisson stream of jobs (rows log-uniform in [1024, 32768], columns in
      [8192, 65536], half case-only, gamma ~ U[0.01, 0.1]) admitted by an HBM budget.
    Traces of the batched runs are compared with the sequential ones (bit for bit).  The
    reference's bit-operation count of the null jobs gives their Tbit-op/s, next to the
    16384 x 65536 single run's (c2_tbitops) as the rate the same kernels reach on a large job."""
    import numpy as np
    from lakcluster_c_amd import gen, lkc
    t0 = time.time()
    pool = gen.make_pool(16384, 16384, 65536, T=3, seed=1)
    print(f"[bench] pool generated in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    jobs = gen.null_jobs(pool, n_null, seed=seed, rows=(2048, 16384), cols=(16384, 65536))
    spec = [(j.rowsA, j.rowsZ, j.colmask, 0.05, 32) for j in jobs]
    rng = np.random.default_rng(seed + 1)
    sjobs = gen.null_jobs(pool, n_serve, seed=seed + 2, rows=(1024, 32768), cols=(8192, 65536))
    sspec = []
    for i, j in enumerate(sjobs):
        g = float(rng.uniform(0.01, 0.1))
        if i % 2:
            sspec.append((np.sort(np.r_[j.rowsA, j.rowsZ]).astype(np.int32), np.zeros(0, np.int32), j.colmask, g, 32))
        else:
            sspec.append((j.rowsA, j.rowsZ, j.colmask, g, 32))
    with lkc.Pool(pool.rows, pool.nrows, pool.ncols, pool.cov, pool.ncov) as P:
        tune = tune_warmup(P, pool, seed)
        t = time.time(); plain = P.batch(spec, max_concurrent=1); tp = time.time() - t
        print(f"[bench] nulls: {n_null} runs one at a time in {tp:.2f} s", file=sys.stderr, flush=True)
        t = time.time(); bat = P.batch(spec, max_concurrent=k_null); tb = time.time() - t
        print(f"[bench] nulls: {n_null} runs {k_null} at a time in {tb:.2f} s", file=sys.stderr, flush=True)
        same = all(a["status"] == 0 and a["trace"].tobytes() == b["trace"].tobytes()
                   and np.array_equal(a["xdrop_a"], b["xdrop_a"]) for a, b in zip(plain, bat))
        # serving: arrivals at about 70% of the batched throughput measured above (scaled
        # by the serving jobs' mean sequential service time relative to the null jobs')
        t = time.time(); P.batch(sspec[:4], max_concurrent=1); ts = (time.time() - t) / 4
        rate = 0.7 * (n_null / tb) * (tp / n_null) / max(ts, 1e-3)
        arr = np.cumsum(rng.exponential(1.0 / rate, size=len(sspec)))
        arr -= arr[0]
        t = time.time(); srv = P.batch(sspec, max_concurrent=k_serve, arrivals=arr); tsrv = time.time() - t
    lkc.tune_set(tune.pop("mode_before"))
    lat = np.array([r["end_s"] - r["arrival_s"] for r in srv])
    ok = all(r["status"] == 0 for r in srv)
    print(f"[bench] serving: {len(sspec)} jobs in {tsrv:.2f} s", file=sys.stderr, flush=True)
    null_ops = float(sum(r["ops"] for r in plain))
    return {
        "batched_nulls": {
            "runs": n_null, "plain_runs_per_hour": n_null / tp * 3600, "batched_runs_per_hour": n_null / tb * 3600,
            "ratio": tp / tb, "concurrent": k_null, "identical_to_plain": bool(same),
            "passes_mean": float(np.mean([r["n_iter"] for r in plain])),
            "bit_ops_per_run_mean": null_ops / n_null,
            "plain_Tbitops": null_ops / tp / 1e12, "batched_Tbitops": null_ops / tb / 1e12,
            "batched_vs_single_run_rate": (null_ops / tb / 1e12) / c2_tbitops if c2_tbitops else None,
            "tuning": tune,
            "sample": f"{n_null} permutation-null jobs (rows log-U[2048,16384] of the pooled "
                      "16384+16384 patients x 65536 SNPs, labels re-shuffled, columns log-U[16384,65536], gamma 0.05, B 32)",
        },
        "serving": {
            "jobs": len(sspec), "concurrent": k_serve, "arrival_rate_per_s": rate, "jobs_per_hour": len(sspec) / tsrv * 3600,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "all_ok": bool(ok),
            "sample": "Poisson arrivals; rows log-U[1024,32768], columns log-U[8192,65536] from the 16384+16384 x 65536 "
                      "pool, half case-only, gamma ~ U[0.01,0.1], B 32 (job sizes capped at the pool)",
        },
    }


def nulls_leg(n_null, k_null, seed=7):
    """The permutation-null jobs of batched_legs alone: n_null jobs one at a time (the
    plain GPU path) and, for k_null > 1, k_null at a time, traces compared."""
    import numpy as np
    from lakcluster_c_amd import gen, lkc
    pool = gen.make_pool(16384, 16384, 65536, T=3, seed=1)
    jobs = gen.null_jobs(pool, n_null, seed=seed, rows=(2048, 16384), cols=(16384, 65536))
    spec = [(j.rowsA, j.rowsZ, j.colmask, 0.05, 32) for j in jobs]
    out = {"runs": n_null}
    with lkc.Pool(pool.rows, pool.nrows, pool.ncols, pool.cov, pool.ncov) as P:
        tune = tune_warmup(P, pool, seed)
        # phase bounds on the monotonic clock (the profiler's kernel timestamps use it too)
        m0 = time.monotonic_ns()
        t = time.time(); plain = P.batch(spec, max_concurrent=1); tp = time.time() - t
        out["phase_ns"] = {"plain": [m0, time.monotonic_ns()]}
        print(f"[bench] nulls: {n_null} runs one at a time in {tp:.2f} s", file=sys.stderr, flush=True)
        null_ops = float(sum(r["ops"] for r in plain))
        out["tuning"] = tune
        out.update({"plain_s": tp, "plain_runs_per_hour": n_null / tp * 3600,
                    "passes_mean": float(np.mean([r["n_iter"] for r in plain])),
                    "bit_ops_per_run_mean": null_ops / n_null, "plain_Tbitops": null_ops / tp / 1e12})
        if k_null > 1:
            m0 = time.monotonic_ns()
            t = time.time(); bat = P.batch(spec, max_concurrent=k_null); tb = time.time() - t
            out["phase_ns"]["batched"] = [m0, time.monotonic_ns()]
            print(f"[bench] nulls: {n_null} runs {k_null} at a time in {tb:.2f} s", file=sys.stderr, flush=True)
            out.update({"concurrent": k_null, "batched_s": tb, "batched_runs_per_hour": n_null / tb * 3600,
                        "ratio": tp / tb, "batched_Tbitops": null_ops / tb / 1e12,
                        "identical_to_plain": all(a["status"] == 0 and a["trace"].tobytes() == b["trace"].tobytes()
                                                  for a, b in zip(plain, bat))})
    lkc.tune_set(tune.pop("mode_before"))
    return out


def dcc_leg(text, max_iter=-1):
    """The dcc half-loop elimination (dexcluster_driver) on the same problem: one untimed
    run, one timed run; its dominant kernel is the a0d1-weighted An_ZtSn_ww contraction."""
    from lakcluster_c_amd import lkc
    dtext = text.replace("GLOBAL_TEST_TYPE= lakcluster_driver;", "GLOBAL_TEST_TYPE= dexcluster_driver;")
    run = lkc.Bcc(dtext)
    run.set_profiling(True)
    run.run(max_iter)
    run.reset()
    t0 = time.perf_counter()
    run.run(max_iter)
    el = time.perf_counter() - t0
    launches, ms, work = run.kernel_stats("ztsn")
    out = {"elim_run_s": el, "passes": run.n_iter, "bit_ops_per_run": run.ops, "value": run.ops / el / 1e12,
           "unit": "Tbit-op/s", "host_quicksorts": run.sort_fallbacks}
    run.close()
    # the contraction streams every active packed row once per launch: HBM-bound
    ach = work / (ms / 1e3) / 1e9 if ms > 0 else None
    out["roofline"] = {"kernel": "An_ZtSn_ww contraction", "bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS,
                       "unit": "GB/s", "frac": ach / PEAK_HBM_GBS if ach else None, "launches": launches,
                       "avg_launch_ms": ms / launches if launches else None}
    print(f"[bench] dcc: {el:.3f} s, {out['passes']} passes, ztsn {launches} launches {ms:.1f} ms",
          file=sys.stderr, flush=True)
    return out


def dcc_bins_leg(in_path, kw, root, nbins=4, ny=8192):
    """The dcc elimination with several bins and Y/W genetic controls: the same patients
    split into nbins bins with ny Y/W columns added, through lkc_dcc_driver_bins (the
    whole call: .b16 files read and uploaded, every pass, the three files written); the
    one-bin split through the same entry beside it.  One untimed call each first."""
    from lakcluster_c_amd import gen, lkc
    ds = gen.Dataset(in_path, os.path.join(root, "unused_out"), kw["nA"], kw["nZ"], kw["ncols"], kw["T"],
                     kw["gamma"], kw["B"])
    out = {"workload": f"{kw['nA']}+{kw['nZ']}x{kw['ncols']} split into {nbins} bins, {ny} Y/W columns",
           "includes": "file load, upload, every pass, output files"}
    for tag, nb, y in (("bins", nbins, ny), ("one_bin", 1, 0)):
        t0 = time.time()
        bds = gen.split_into_bins(ds, root, f"dccbins_{tag}", nb, ny=y, seed=7)
        text = open(bds.in_path).read()
        gen_s = time.time() - t0
        odir = os.path.join(root, f"dccbins_{tag}_out")
        lkc.dcc_driver_bins(text, odir)
        t0 = time.perf_counter()
        lkc.dcc_driver_bins(text, odir)
        el = time.perf_counter() - t0
        passes = sum(1 for _ in open(os.path.join(odir, "out_trace.txt")))
        out[tag] = {"call_s": el, "passes": passes, "bins": nb, "Y_cols": y}
        print(f"[bench] dcc {tag}: {el:.3f} s, {passes} passes (split in {gen_s:.1f} s)", file=sys.stderr, flush=True)
    return out


def large_leg(rows=65536, cols=524288, passes=20, gamma=0.01, B=24, seed=1, check_valu=False, repeat=False):
    """The large single run (A and Z rows x cols each, p_j ~ U[0.01, 0.5], biclusters
    of rows**0.5, rows**0.55 and rows**0.6 rows planted in A, T/S 3 columns, gamma 0.01,
    B_MLT 24): the initial loop-subscores and the first ``passes`` elimination passes,
    each timed, inputs drawn on the device and resident in HBM.  With check_valu the
    same passes are repeated with the popcount (VALU) kernels, whose indexing is
    separate code, and the traces must agree bit for bit."""
    import numpy as np
    from lakcluster_c_amd import gen
    t0 = time.time()
    pool = gen.make_pool_random(rows, rows, cols, T=3, pmin=0.01, pmax=0.5, planted=(0.5, 0.55, 0.6), seed=seed)
    planted_rows = pool.planted_rows
    print(f"[bench] large: pool {rows}+{rows} x {cols} drawn in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    out = {"workload": f"large_{rows}x{cols}", "rows_A": rows, "rows_Z": rows, "cols": cols, "covariates": 3,
           "gamma": gamma, "B_MLT": B}
    with pool.pool as P:
        ra = np.arange(rows, dtype=np.int32); rz = np.arange(rows, 2 * rows, dtype=np.int32)
        traces = []
        kerns = ["mfma", "valu"] if check_valu else ["mfma", "mfma"] if repeat else ["mfma"]
        xdrops = []
        for rep, kern in enumerate(kerns):
            t0 = time.time()
            run = P.job(ra, rz, None, gamma, B)
            run.set_profiling(True)
            if kern != "mfma":
                run.set_kernels(kern)
            t_load = time.time() - t0
            t0 = time.perf_counter()
            run.init_lf()
            t_init = time.perf_counter() - t0
            ops_init = run.ops
            print(f"[bench] large ({kern}): loaded in {t_load:.1f} s, initial subscores {t_init:.2f} s, "
                  f"{ops_init:.3e} bit-ops", file=sys.stderr, flush=True)
            t0 = time.perf_counter(); tl = t0
            for k in range(passes):
                if run.iterate():
                    break
                if time.perf_counter() - tl > 20 or k + 1 == passes:
                    tl = time.perf_counter()
                    print(f"[bench] large ({kern}): {run.n_iter} passes, {tl - t0:.2f} s", file=sys.stderr, flush=True)
            t_loop = time.perf_counter() - t0
            tr = run.trace()
            traces.append(tr)
            xdrops.append(run.xdrop())
            if kern == "mfma" and rep == 0:
                launches, ms, work = run.kernel_stats("ww")
                ach = work / (ms / 1e3) / 1e12 if ms > 0 else None
                out.update({
                    "load_s": t_load, "init_s": t_init, "passes": run.n_iter, "passes_s": t_loop,
                    "bit_ops_init": ops_init, "bit_ops": run.ops,
                    "value": run.ops / (t_init + t_loop) / 1e12, "unit": "Tbit-op/s",
                    "trace_finite": bool(np.isfinite(tr).all()),
                    "finished": bool(run.n_iter > 0 and tr[-1, 1] <= 2),
                    "rows_cols_after": [int(tr[-1, 1]), int(tr[-1, 2])] if len(tr) else None,
                    "roofline": {"kernel": "At_T_YnWt_S_Zn_ww contraction", "bound": "mfma", "achieved": ach,
                                 "peak": PEAK_INT8_TOPS, "unit": "TOP/s",
                                 "frac": ach / PEAK_INT8_TOPS if ach else None, "launches": launches,
                                 "avg_launch_ms": ms / launches if launches else None},
                })
            if kern == "mfma" and rep == 0 and out["finished"]:
                # the planted biclusters should be the last A rows eliminated
                xa, _ = run.xdrop()
                order = xa[xa[:, 0] >= 0, 0]
                planted = np.unique(np.concatenate(planted_rows))
                for m in sorted(len(r) for r in planted_rows):
                    out.setdefault("planted_recovery", []).append(
                        {"last_rows": m, "planted_frac": float(np.isin(order[-m:], planted).mean())})
            run.close()
        if check_valu:
            out["valu_trace_identical"] = bool(traces[0].tobytes() == traces[1].tobytes())
        if repeat and not check_valu:
            # a second whole run on a fresh handle: the trace (bit patterns) and the xdrop
            # lists must repeat exactly
            out["repeat_identical"] = bool(traces[0].tobytes() == traces[1].tobytes()
                                           and all(np.array_equal(a, b) for a, b in zip(xdrops[0], xdrops[1])))
    return out


def serve_large_leg(n_jobs=24, k=8, seed=11, load=0.7):
    """Serving at the full job-size range: a device-drawn pool of 65536 cases + 65536
    controls x 524288 SNPs (p_j ~ U[0.05, 0.5], a planted bicluster in A, T/S 3 columns);
    jobs draw rows log-uniform in [1024, 65536] and a column subset log-uniform in
    [8192, 524288] (a column mask), half case-only, gamma ~ U[0.01, 0.1], B 32.  The jobs
    first run as one closed batch (all present at time 0, k at a time, admitted by the
    HBM budget) for the capacity in jobs/hour, then again as a Poisson stream at
    ``load`` times that capacity for p50/p99 latency; every job's trace must be the same
    in both runs."""
    import numpy as np
    from lakcluster_c_amd import gen
    t0 = time.time()
    pool = gen.make_pool_random(65536, 65536, 524288, T=3, pmin=0.05, pmax=0.5, planted=(0.55,), seed=seed)
    print(f"[bench] serve-large: pool drawn in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    rng = np.random.default_rng(seed + 1)
    sjobs = gen.null_jobs(pool, n_jobs, seed=seed + 2, rows=(1024, 65536), cols=(8192, 524288))
    spec, sizes = [], []
    for i, j in enumerate(sjobs):
        g = float(rng.uniform(0.01, 0.1))
        if i % 2:
            spec.append((np.sort(np.r_[j.rowsA, j.rowsZ]).astype(np.int32), np.zeros(0, np.int32), j.colmask, g, 32))
        else:
            spec.append((j.rowsA, j.rowsZ, j.colmask, g, 32))
        sizes.append((int(spec[-1][0].size), int(spec[-1][1].size), int(j.ncols_active)))
    with pool.pool as P:
        t = time.time(); closed = P.batch(spec, max_concurrent=k); tc = time.time() - t
        print(f"[bench] serve-large: {n_jobs} jobs as one batch in {tc:.2f} s", file=sys.stderr, flush=True)
        rate = load * n_jobs / tc
        arr = np.cumsum(rng.exponential(1.0 / rate, size=n_jobs))
        arr -= arr[0]
        t = time.time(); srv = P.batch(spec, max_concurrent=k, arrivals=arr); ts = time.time() - t
        print(f"[bench] serve-large: {n_jobs} Poisson arrivals in {ts:.2f} s", file=sys.stderr, flush=True)
    lat = np.array([r["end_s"] - r["arrival_s"] for r in srv])
    svc = np.array([r["end_s"] - r["start_s"] for r in srv])
    same = all(a["status"] == 0 and b["status"] == 0 and a["trace"].tobytes() == b["trace"].tobytes()
               for a, b in zip(closed, srv))
    return {
        "jobs": n_jobs, "concurrent": k, "capacity_jobs_per_hour": n_jobs / tc * 3600,
        "arrival_rate_per_s": rate, "jobs_per_hour": n_jobs / ts * 3600,
        "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
        "service_p50_s": float(np.percentile(svc, 50)), "service_max_s": float(svc.max()),
        "identical_batch_and_stream": bool(same),
        "job_sizes": sizes,
        "sample": f"{n_jobs} jobs: rows log-U[1024,65536] of a 65536+65536 x 524288 pool, column masks log-U"
                  "[8192,524288], half case-only, gamma ~ U[0.01,0.1], B 32; Poisson arrivals at "
                  f"{load:.0%} of the closed-batch capacity",
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="case_control_16384x65536")
    ap.add_argument("--max-iter", type=int, default=-1, help="elimination passes per run (-1: to the end)")
    ap.add_argument("--data-root", default=os.path.join(tempfile.gettempdir(), "lkc_bench_data"))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-null and serving legs")
    ap.add_argument("--no-dcc", action="store_true", help="skip the dcc half-loop elimination leg")
    ap.add_argument("--dcc-only", action="store_true", help="run only the dcc leg and print its JSON object")
    ap.add_argument("--dcc-bins-only", action="store_true",
                    help="run only the several-bin dcc leg (4 bins, 8192 Y/W columns) and print its JSON object")
    ap.add_argument("--large", type=int, nargs=2, metavar=("ROWS", "COLS"),
                    help="run only the large single-run leg (e.g. 65536 524288) for --max-iter passes "
                         "(default: to the end) and print its JSON object")
    ap.add_argument("--serve-large", type=int, metavar="JOBS",
                    help="run only the serving leg at the full job-size range with JOBS jobs and print its JSON object")
    ap.add_argument("--nulls", type=int, metavar="RUNS",
                    help="run only the permutation-null leg with RUNS jobs (plain, then --concurrent at a time)")
    ap.add_argument("--concurrent", type=int, default=8, help="with --nulls: jobs at a time (1: plain only)")
    ap.add_argument("--null-runs", type=int, default=256,
                    help="permutation-null jobs of the batched leg (the 256-run batch by default)")
    ap.add_argument("--kernels", choices=["mfma", "valu"], default="mfma",
                    help="contraction kernels of the main leg: int8 matrix cores (default) or VALU popcount")
    ap.add_argument("--repeat", action="store_true",
                    help="with --large: run the elimination a second time on a fresh handle and compare")
    ap.add_argument("--check-valu", action="store_true",
                    help="with --large: repeat the passes with the popcount kernels and compare traces")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 1 GiB buffer three times (the reference kernel for scaling HBM counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 1 << 28   # 1 GiB of float32
        x = torch.ones(n, device="cuda"); y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 4 * n, "copies": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this path runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.nulls:
        print(json.dumps({"batched_nulls": nulls_leg(args.nulls, args.concurrent)}), flush=True)
        return
    if args.serve_large:
        print(json.dumps({"serve_large": serve_large_leg(args.serve_large)}), flush=True)
        return
    if args.large:
        passes = args.max_iter if args.max_iter > 0 else 1 << 30
        print(json.dumps({"large": large_leg(args.large[0], args.large[1], passes, check_valu=args.check_valu,
                                               repeat=args.repeat)}),
              flush=True)
        return
    from lakcluster_c_amd import gen, lkc
    os.makedirs(args.data_root, exist_ok=True)
    kw = dict(gen.CONFIGS[args.config])
    kw.setdefault("seed", 1)
    print(f"[bench] generating {args.config} ...", file=sys.stderr, flush=True)
    in_path = make_data(args.data_root, args.config, kw)
    text = open(in_path).read()
    if args.dcc_only:
        print(json.dumps({"workload": args.config, "dcc": dcc_leg(text, args.max_iter)}), flush=True)
        return
    if args.dcc_bins_only:
        print(json.dumps({"workload": args.config, "dcc_bins": dcc_bins_leg(in_path, kw, args.data_root)}), flush=True)
        return
    print("[bench] loading ...", file=sys.stderr, flush=True)
    run = lkc.Bcc(text)
    run.set_profiling(True)
    if args.kernels != "mfma":
        run.set_kernels(args.kernels)
    for w in range(args.warmup):
        t0 = time.perf_counter()
        run.run(args.max_iter)
        print(f"[bench] warmup {w}: {time.perf_counter() - t0:.2f} s, {run.n_iter} passes", file=sys.stderr, flush=True)
        run.reset()
    elim, ops, kms, kwork, kl = [], [], 0.0, 0.0, 0
    fallbacks = 0
    qr_l, qr_ms, qr_work = 0, 0.0, 0.0
    n_iter = 0
    step_ns = []       # the timed steps on the monotonic clock (profiler timestamps use it too)
    for k in range(args.steps):
        m0 = time.monotonic_ns()
        t0 = time.perf_counter()
        run.run(args.max_iter)
        t1 = time.perf_counter()
        step_ns.append([m0, time.monotonic_ns()])
        elim.append(t1 - t0)
        ops.append(run.ops)
        n_iter = run.n_iter
        fallbacks = run.sort_fallbacks
        launches, ms, work = run.kernel_stats("ww")
        kl += launches; kms += ms; kwork += work
        ql, qms, qwork = run.kernel_stats("qr")
        qr_l += ql; qr_ms += qms; qr_work += qwork
        print(f"[bench] step {k}: {t1 - t0:.3f} s, {run.n_iter} passes, {run.ops:.3e} bit-ops, "
              f"ww {launches} launches {ms:.1f} ms", file=sys.stderr, flush=True)
        if k + 1 < args.steps:
            run.reset()
    shape = run.shape()
    run.close()
    tot_s = sum(elim)
    value = sum(ops) / tot_s / 1e12
    achieved = kwork / (kms / 1e3) / 1e12 if kms > 0 else None
    line = {
        "metric": "elim-run s & Tbitop/s at 16k×64k; batched null runs/hr vs plain, 1 GPU",
        "value": value,
        "unit": "Tbit-op/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * tot_s / args.steps,
        "elim_run_s": tot_s / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8/int32 popcount, f64",
        "data": "synthetic (seeded generator: Bernoulli columns, planted rank-1 bicluster)",
        "config": {"workload": args.config, "kernels": args.kernels, "rows_A": shape[0], "rows_Z": shape[1], "cols": shape[2],
                   "covariates": shape[3], "gamma": kw["gamma"], "B_MLT": kw["B"],
                   "passes": n_iter, "bit_ops_per_run": ops[-1] if ops else None,
                   "host_quicksorts": fallbacks if args.steps else None},
        "roofline": {
            "kernel": "At_T_YnWt_S_Zn_ww contraction",
            "bound": "mfma",
            "achieved": achieved,
            "peak": PEAK_INT8_TOPS,
            "unit": "TOP/s",
            "frac": achieved / PEAK_INT8_TOPS if achieved else None,
            "traffic": _traffic(args.config),
            "launches": kl,
            "avg_launch_ms": kms / kl if kl else None,
        },
    }
    line["step_ns"] = step_ns
    # the QR row sums (AnZt_S_WnYt_vv) stream the loop-subscores: HBM-bound, 8 bytes per
    # (row, column) pair read, HIP events around each launch on the engine's stream
    qach = qr_work / (qr_ms / 1e3) / 1e9 if qr_ms > 0 else None
    line["qr_roofline"] = {"kernel": "AnZt_S_WnYt_vv row sums", "bound": "hbm", "achieved": qach,
                           "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": qach / PEAK_HBM_GBS if qach else None,
                           "bytes_per_launch": qr_work / qr_l if qr_l else None, "launches": qr_l,
                           "avg_launch_ms": qr_ms / qr_l if qr_l else None}
    if not args.no_dcc:
        print("[bench] dcc half-loop elimination ...", file=sys.stderr, flush=True)
        line["dcc"] = dcc_leg(text, args.max_iter)
        if args.max_iter < 0 and kw.get("nZ", 0) > 0:
            print("[bench] dcc with several bins and Y/W controls ...", file=sys.stderr, flush=True)
            line["dcc_bins"] = dcc_bins_leg(in_path, kw, args.data_root)
    if not args.no_batch:
        print("[bench] batched nulls and serving ...", file=sys.stderr, flush=True)
        line.update(batched_legs(n_null=args.null_runs, c2_tbitops=value))
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(args.data_root)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
