#!/usr/bin/env python3
"""Headline benchmark: brute-force L2 kNN QPS on one MI355X.

Workload (the metric's configuration): database 10,000,000 x 96 unit-normalised
standard-Gaussian vectors (generator seed 4), 10,000 queries from the same
distribution (seed 5), k = 100. A step is one IndexBF.search of the whole
10,000-query batch with queries and outputs resident in HBM.

Forms measured on the same resident index:
  * bf16  (headline `value`): BF16 MFMA candidates, over-fetch k' = 4k = 400,
    exact FP32 re-rank, FP64 re-rank of the boundary band;
  * fp32  (plain path, the drop-in): FP32 MFMA candidates + canonical FP32
    re-rank (bit-exact to the canonical restatement);
  * fp8   : e4m3 + per-row power-of-two scale candidates, same re-rank.
recall@100 and the max relative distance error are measured against the FP64
oracle on a query subset; the plain path is checked bit-exact against the
canonical FP32 restatement on that subset.

roofline: the dominant kernel is the full-database filter launch (the fused
contraction). Its algorithmic work is 2 * nq * rows * d flops per launch, rows =
the database rows it scores (all n, less the tiles of a carried last sample
level, whose survivors it reuses: stats rows_filter_full); its
duration is measured live with HIP events recorded on the search stream around
that launch. cpu_baseline: the canonical FP32 restatement of the reference's
searchL2Distance, run on the host cores over a bounded query sample.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

MI355X_PEAK = {"bf16": 2500.0, "fp8": 5000.0, "fp32": 157.3}  # dense TFLOP/s
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--d", type=int, default=96)
    ap.add_argument("--k", type=int, default=100)
    ap.add_argument("--kprime-factor", type=int, default=4,
                    help="k'/k of the BF16/FP8 forms (the metric's setting: 4)")
    ap.add_argument("--data", default="gaussian", choices=["gaussian", "uniform01", "mixture", "mixture1024"],
                    help="headline data: unit Gaussian (seeds 4/5, the metric's), U[0,1) (seeds 0/1), "
                         "256 centres U[-1,1) + N(0, 0.1^2) (seed 2, queries seed 30), or 1024 centres "
                         "+ N(0, 0.05^2) with 5%% exact duplicate rows (seed 6, queries seed 7)")
    ap.add_argument("--precision", default="bf16", choices=["bf16", "fp8", "fp32"])
    ap.add_argument("--residual-clusters", type=int, default=0,
                    help="BF16/FP8: centroids of the cluster-residual copy (0: off)")
    ap.add_argument("--forms", default="bf16,fp32,fp8",
                    help="forms to time (the --precision form is the headline value)")
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--recall-queries", type=int, default=1000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (grid, survivor target, filter variant; k' stays 4k)")
    ap.add_argument("--mode", default="headline",
                    choices=["headline", "serving", "copy", "pmc-pass", "rplsh", "smallbatch"],
                    help="headline: the metric's workload; serving: the 64-index request loop; "
                         "copy: device-to-device copies of --copy-gib GiB (HBM bandwidth and the "
                         "reference for scaling PMC byte counters); pmc-pass: one search of the "
                         "headline form (what the traffic passes profile); smallbatch: nq in "
                         "--sb-nq against 2M-row indexes, the database stream's GB/s (HBM-bound regime)")
    ap.add_argument("--sb-shapes", default="2000000x256,2000000x64", help="smallbatch: NxD,... indexes")
    ap.add_argument("--sb-nq", default="1,8,64,512", help="smallbatch: batch sizes")
    ap.add_argument("--sb-forms", default="bf16,fp32", help="smallbatch: precisions")
    ap.add_argument("--sb-steps", type=int, default=30, help="smallbatch: timed searches per point")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 --pmc passes that measure the filter's HBM traffic")
    ap.add_argument("--tuning", default="", help="pmc-pass: JSON tuning entry to install")
    ap.add_argument("--copy-gib", type=float, default=1.0)
    ap.add_argument("--lsh-data", default="c2,uniform16", help="rplsh: workloads (c2, uniform16)")
    ap.add_argument("--lsh-n", type=int, default=1_000_000, help="rplsh: rows of the uniform16 workload")
    ap.add_argument("--lsh-configs", default="",
                    help="rplsh: tables:bits:probes:budget[:prerank],... (default: built-in list)")
    ap.add_argument("--rate", type=float, default=400.0, help="serving: Poisson arrival rate (req/s)")
    ap.add_argument("--requests", type=int, default=4096, help="serving: number of requests")
    ap.add_argument("--indexes", type=int, default=64, help="serving: resident indexes")
    ap.add_argument("--slots", type=int, default=4, help="serving: concurrent batches")
    ap.add_argument("--persistent-max-queries", type=int, default=256,
                    help="serving: batches up to this many queries run in the persistent kernel (0: off)")
    ap.add_argument("--serving-precision", default="bf16", choices=["bf16", "fp8", "fp32"])
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): this build has no "
              "multi-GPU path; the per-GPU value is the N=1 line.", file=sys.stderr)
        return 2

    import curplsh_amd as ca

    if ca.device_count() < 1:
        log("no GPU visible")
        return 1
    if args.mode == "serving":
        return serving(args, ca)
    if args.mode == "copy":
        return copy_probe(args, ca)
    if args.mode == "pmc-pass":
        return pmc_pass(args, ca)
    if args.mode == "rplsh":
        return rplsh_front_end(args, ca)
    if args.mode == "smallbatch":
        return small_batch(args, ca)
    n, nq, d, k = args.n, args.nq, args.d, args.k
    log(f"[bench] generating {n}x{d} database and {nq} queries")
    t0 = time.time()
    x, q = headline_data(ca, args.data, n, nq, d)
    log(f"[bench] generated in {time.time() - t0:.1f}s")

    res = ca.DeviceResources(0)
    forms = [f for f in args.forms.split(",") if f]
    if args.precision not in forms:
        forms.insert(0, args.precision)
    forms.sort(key=lambda f: f != args.precision)

    idx = ca.IndexBF(res, d, ca.IndexBFConfig(precision=args.precision, kprime_factor=args.kprime_factor,
                                              residual_clusters=args.residual_clusters))
    t0 = time.time()
    idx.add(x)
    log(f"[bench] index built (add incl. H2D, norms, encode) in {time.time() - t0:.1f}s")
    dq = ca.DeviceArray.from_host(q)
    dI = ca.DeviceArray((nq, k), np.int32)
    dD = ca.DeviceArray((nq, k), np.float32)

    results = {}
    tunings = {}
    for form in forms:
        idx.set_precision(form, args.kprime_factor)
        if not args.no_tune:
            # Untimed: tunes grid size, survivor target and filter variant for
            # this batch; target_recall 0 keeps the set k'/k (the metric's: 4).
            t0 = time.time()
            tunings[form] = idx.autotune(q[:256], k, nq, 0.0)
            log(f"[bench] {form} autotune in {time.time() - t0:.1f}s: {tunings[form]}")
        steps = args.steps if form == args.precision else max(1, args.plain_steps)
        for _ in range(args.warmup):
            idx.search(dq, k, dI, dD)
        ca.call("rplsh_device_sync")
        filt, tot, sel, rr, samp = [], [], [], [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            idx.search(dq, k, dI, dD)
            s = idx.last_stats()
            filt.append(s["ms_filter_full"])
            tot.append(s["ms_total"])
            sel.append(s["ms_select"])
            rr.append(s["ms_rerank"])
            samp.append(s["ms_filter_samples"])
        ca.call("rplsh_device_sync")
        wall = time.perf_counter() - t0
        ms = wall * 1000.0 / steps
        fms = float(np.mean(filt))
        st = idx.last_stats()
        # rows the full launch scores: a carried last sample level's tiles are skipped
        rows = int(st["rows_filter_full"])
        flops = 2.0 * nq * rows * d
        results[form] = {
            "qps": nq / (ms / 1000.0),
            "ms_per_step": ms,
            "ms_filter_full": fms,
            "ms_filter_samples": float(np.mean(samp)),
            "ms_select": float(np.mean(sel)),
            "ms_rerank": float(np.mean(rr)),
            "filter_tflops": flops / (fms / 1000.0) / 1e12,
            "rows_filter_full": rows,
            "batch_filter_tflops": 2.0 * nq * n * d / ((fms + float(np.mean(samp))) / 1000.0) / 1e12,
            "steps": steps,
            "stats": st,
            "I": dI.to_host(),
            "D": dD.to_host(),
        }
        log(f"[bench] {form}: {ms:.2f} ms/step, QPS {results[form]['qps']:.0f}, filter {fms:.2f} ms "
            f"({results[form]['filter_tflops']:.1f} TFLOP/s), samples {np.mean(samp):.2f} ms, "
            f"select {np.mean(sel):.2f} ms, rerank {np.mean(rr):.2f} ms, stats {st}")

    head = results[args.precision]
    extra = {}
    # Accuracy against the FP64 oracle / the canonical restatement on a subset.
    R = min(args.recall_queries, nq)
    oracle_mod = None
    if R > 0:
        import oracle as oracle_mod  # the checker, not the measured path

        t0 = time.time()
        I64, _ = oracle_mod.search_fp64(x, q[:R], k)
        log(f"[bench] FP64 oracle on {R} queries in {time.time() - t0:.1f}s")
        for form, r in results.items():
            rec = oracle_mod.recall(r["I"][:R], I64)
            dp = oracle_mod.pair_dist_fp64(x, q[:R], r["I"][:R])
            err = float(np.max(np.abs(r["D"][:R].astype(np.float64) - dp) / np.maximum(dp, 1e-30)))
            extra[f"{form}_recall_at_{k}"] = rec
            extra[f"{form}_max_rel_dist_err"] = err
        if "fp32" in results:
            t0 = time.time()
            Ic, Dc = oracle_mod.search_canonical(x, q[:R], k)
            extra["fp32_bitexact_vs_canonical"] = bool(
                np.array_equal(Ic, results["fp32"]["I"][:R])
                and np.array_equal(Dc.view(np.uint32), results["fp32"]["D"][:R].view(np.uint32)))
            log(f"[bench] canonical check on {R} queries in {time.time() - t0:.1f}s")

    # CPU baseline: canonical FP32 restatement (port) on a bounded query sample.
    cpu = None
    if args.cpu_seconds > 0:
        if oracle_mod is None:
            import oracle as oracle_mod
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)) or 1)
        # Sample sized so one run takes ~cpu_seconds / 3 (the sizing runs are the
        # warm-up); the reported time is the median of three runs.
        S = min(nq, threads)
        for _ in range(5):
            t0 = time.perf_counter()
            oracle_mod.search_canonical(x, q[:S], k, threads=threads)
            tc = time.perf_counter() - t0
            if tc >= args.cpu_seconds / 6.0 or S >= nq:
                break
            S = int(min(nq, max(2 * S, S * args.cpu_seconds / 3.0 / max(tc, 1e-6))))
            S = max(threads, (S // threads) * threads)
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            oracle_mod.search_canonical(x, q[:S], k, threads=threads)
            runs.append(time.perf_counter() - t0)
        tc = float(np.median(runs))
        model = ""
        try:
            with open("/proc/cpuinfo") as fh:
                model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), "")
        except OSError:
            pass
        cpu = {"value": S / tc, "unit": "QPS", "cores": threads, "kind": "port",
               "gflops": round(2.0 * S * n * d / tc / 1e9, 1), "cpu_model": model,
               "affinity_cpus": len(os.sched_getaffinity(0)), "host_cpus": os.cpu_count(),
               "runs_s": [round(t, 3) for t in runs],
               "sample": f"first {S} of the {nq} queries against the full {n}x{d} database, k={k}, "
                         f"canonical FP32 restatement of searchL2Distance (oracle/oracle.c), "
                         f"{threads} OpenMP threads, median of 3 runs {tc:.1f}s each"}
        log(f"[bench] CPU baseline {cpu['value']:.1f} QPS on {threads} threads ({S} queries, {tc:.1f}s)")

    # Measured ceilings of this device: MFMA (random and constant operands) and
    # a device-to-device copy.
    measured = {"mfma_tflops_random": {}, "mfma_tflops_constant": {}}
    for f in ("bf16", "fp8", "fp32"):
        measured["mfma_tflops_random"][f] = round(mfma_peak(f, True), 1)
        measured["mfma_tflops_constant"][f] = round(mfma_peak(f, False), 1)
    measured["hbm_copy_gbs"] = round(copy_bandwidth(ca, 1 << 30, 3), 1)
    log(f"[bench] measured peaks {measured}")

    form = args.precision
    peak = MI355X_PEAK[form]
    ach = head["filter_tflops"]
    variant = tunings[form]["variant"] if form in tunings else 0
    traffic = {}
    if not args.no_traffic:
        traffic = pmc_traffic(args, tunings.get(form))
        log(f"[bench] traffic: {traffic}")
    line = {
        "metric": "QPS at 10M x 96, k=100: BF16/FP8+rerank vs FP32 plain, recall@100 >= 0.999",
        "value": round(head["qps"], 1),
        "unit": "QPS",
        "n_gpus": 1,
        "steps": head["steps"],
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": form,
        "data": DATA_DESC[args.data],
        "config": {
            "workload": f"{n}x{d} database ({args.data}), {nq} queries, k={k}, {form} candidates k'={args.kprime_factor}k + FP32 re-rank"
                        + (" + FP64 boundary" if form != "fp32" else ""),
            "n": n, "d": d, "nq": nq, "k": k, "precision": form, "kprime": args.kprime_factor * k if form != "fp32" else None,
            "residual_clusters": args.residual_clusters,
        },
        "roofline": {
            "bound": "mfma",
            "achieved": round(ach, 2),
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": round(ach / peak, 4),
            "traffic": traffic.get("bytes"),
            # the tuned variant's low byte is the full level's kernel (bits 8-15: the sample levels')
            "kernel": f"{kernel_of(variant, args.residual_clusters > 0)}<{form}> full-database launch",
            "algorithmic_flops_per_launch": 2.0 * nq * head["rows_filter_full"] * d,
            "rows_per_launch": head["rows_filter_full"],
            "frac_of_measured_mfma": round(ach / measured["mfma_tflops_random"][form], 4),
            "ms_per_launch": round(head["ms_filter_full"], 4),
        },
        "cpu_baseline": cpu,
        "measured_peaks": measured,
    }
    for f, r in results.items():
        line[f"{f}_qps"] = round(r["qps"], 1)
        line[f"{f}_ms_per_step"] = round(r["ms_per_step"], 3)
        line[f"{f}_filter_tflops"] = round(r["filter_tflops"], 2)
        line[f"{f}_filter_frac_of_peak"] = round(r["filter_tflops"] / MI355X_PEAK[f], 4)
        # every filter launch of a batch (sample levels + full) against 2 nq n d
        line[f"{f}_batch_filter_tflops"] = round(r["batch_filter_tflops"], 2)
        line[f"{f}_rows_filter_full"] = r["rows_filter_full"]
        line[f"{f}_phase_ms"] = {"filter_full": round(r["ms_filter_full"], 3),
                                 "filter_samples": round(r["ms_filter_samples"], 3),
                                 "select": round(r["ms_select"], 3), "rerank": round(r["ms_rerank"], 3)}
        line[f"{f}_retried_queries"] = r["stats"]["retried_queries"]
        if f in tunings:
            line[f"{f}_tuning"] = {key: tunings[f][key] for key in
                                   ("target_wgs", "c_factor", "rho", "variant", "ms_default", "ms_tuned")}
    if "fp32" in results and form != "fp32":
        line["speedup_vs_plain_fp32"] = round(head["qps"] / results["fp32"]["qps"], 3)
    if traffic:
        line["traffic_detail"] = traffic
    line.update(extra)
    print(json.dumps(line), flush=True)
    return 0


KERNEL_OF_VARIANT = {0: "filter_lds_kernel(wide)", 1: "filter_lds_kernel(narrow)",
                     2: "filter_packed_kernel", 3: "filter_ring_kernel(wide)",
                     4: "filter_ring_kernel(narrow)", 5: "filter_lds_kernel(wide, 8 waves)"}


def kernel_of(variant, residual):
    """Kernel of the full level for a tuned variant. Residual levels run the
    LDS kernel over the residual copy: narrow for variant 1, wide otherwise."""
    if residual:
        return "filter_lds_kernel(residual, narrow)" if variant & 0xFF == 1 else "filter_lds_kernel(residual, wide)"
    return KERNEL_OF_VARIANT.get(variant & 0xFF, "filter")


def pmc_pass(args, ca) -> int:
    """One search of the headline workload with the given tuning: the program
    the traffic passes run under rocprofv3 --pmc."""
    n, nq, d, k = args.n, args.nq, args.d, args.k
    x, q = headline_data(ca, args.data, n, nq, d)
    res = ca.DeviceResources(0)
    idx = ca.IndexBF(res, d, ca.IndexBFConfig(precision=args.precision, kprime_factor=args.kprime_factor))
    idx.add(x)
    if args.tuning:
        idx.set_tuning(json.loads(args.tuning))
    dq = ca.DeviceArray.from_host(q)
    dI = ca.DeviceArray((nq, k), np.int32)
    dD = ca.DeviceArray((nq, k), np.float32)
    idx.search(dq, k, dI, dD)
    ca.call("rplsh_device_sync")
    print(json.dumps({"pmc_pass": "done", "stats": idx.last_stats()}), flush=True)
    return 0


def _pmc_values(outdir, counter):
    """{dispatch id: (kernel name, duration ns, summed counter value)}."""
    import csv
    import glob

    vals = {}
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as fh:
            for r in csv.DictReader(fh):
                if r["Counter_Name"] != counter:
                    continue
                key = r["Dispatch_Id"]
                dur = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
                name, _, v = vals.get(key, (r["Kernel_Name"], dur, 0.0))
                vals[key] = (name, dur, v + float(r["Counter_Value"]))
    return vals


def pmc_traffic(args, tuning) -> dict:
    """HBM bytes of the full-level filter launch: FETCH_SIZE and WRITE_SIZE
    (KiB per dispatch) in two rocprofv3 --pmc passes over one search, each
    scaled by what the same counter reads for a device copy of known size."""
    import shutil
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3")
    if prof is None:
        return {"bytes": None, "note": "rocprofv3 not found"}
    env = dict(os.environ, TMPDIR="/tmp")
    me = os.path.abspath(__file__)
    tune = json.dumps({key: tuning[key] for key in tuning}) if tuning else ""
    progs = {
        "filter": [sys.executable, me, "--mode", "pmc-pass", "--precision", args.precision,
                   "--n", str(args.n), "--nq", str(args.nq), "--d", str(args.d), "--k", str(args.k),
                   "--data", args.data, "--kprime-factor", str(args.kprime_factor)]
                  + (["--tuning", tune] if tune else []),
        "copy": [sys.executable, me, "--mode", "copy", "--copy-gib", "1", "--steps", "2"],
    }
    got = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        for what, prog in progs.items():
            out = tempfile.mkdtemp(prefix="rplsh_pmc_", dir="/tmp")
            cmd = ["timeout", "-s", "KILL", "180", prof, "--pmc", counter, "--output-format", "csv",
                   "-d", out, "-o", "p", "--"] + prog
            log(f"[bench] pmc pass {counter} {what}")
            r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE)
            vals = _pmc_values(out, counter)
            shutil.rmtree(out, ignore_errors=True)
            if r.returncode != 0 or not vals:
                return {"bytes": None, "note": f"rocprofv3 pass {counter}/{what} failed (rc {r.returncode})"}
            if what == "filter":
                # The full-level launch is the longest filter dispatch.
                cands = [v for v in vals.values() if "filter" in v[0]]
                got[(counter, what)] = max(cands, key=lambda v: v[1])[2] if cands else None
            else:
                got[(counter, what)] = max(vals.values(), key=lambda v: v[2])[2]
    copy_bytes = float(1 << 30)
    f_scale = copy_bytes / (got[("FETCH_SIZE", "copy")] * 1024.0)
    w_scale = copy_bytes / (got[("WRITE_SIZE", "copy")] * 1024.0)
    fetch = got[("FETCH_SIZE", "filter")] * 1024.0 * f_scale
    write = got[("WRITE_SIZE", "filter")] * 1024.0 * w_scale
    return {"bytes": round(fetch + write), "fetch_bytes": round(fetch), "write_bytes": round(write),
            "fetch_kib_raw": got[("FETCH_SIZE", "filter")], "write_kib_raw": got[("WRITE_SIZE", "filter")],
            "fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4),
            "calibration": "hipMemcpy device-to-device of 1 GiB: scale = 1 GiB / counter bytes"}


def small_batch(args, ca) -> int:
    """The HBM-bound regime: a few queries against a 2M-row index, where the
    full-level filter's time is the database stream. For each index shape,
    precision and batch size: median device time of IndexBF.search, of its
    full-level filter launch (HIP events on the search stream), and the bytes
    that launch must read -- the candidate copy of the rows it scores plus
    their norms (BF16: 32-row packed tiles of dpad elements; FP32: the
    row-major rows) -- as GB/s against the 8 TB/s HBM peak. BF16/FP8 run both
    paths: the launched pipeline and the persistent kernel (one launch per
    step; its full level is one step, timed the same way)."""
    res = ca.DeviceResources(0)
    shapes = [tuple(int(v) for v in sh.split("x")) for sh in args.sb_shapes.split(",") if sh]
    nqs = [int(v) for v in args.sb_nq.split(",") if v]
    forms = [f for f in args.sb_forms.split(",") if f]
    k = 10
    for n, d in shapes:
        log(f"[smallbatch] {n}x{d} U[-1,1)")
        x = ca.gen_uniform(n, d, 9000 + d, -1.0, 1.0)
        idx = ca.IndexBF(res, d, ca.IndexBFConfig(precision=forms[0]))
        idx.add(x)
        del x
        points = []
        for form in forms:
            idx.set_precision(form, 4)
            elem = {"bf16": 2, "fp8": 1, "fp32": 4}[form]
            if form == "bf16":
                dpad = next(s_ for s_ in (32, 64, 96, 128, 160, 192, 256) if d <= s_)
            elif form == "fp8":
                dpad = next(s_ for s_ in (64, 128, 256) if d <= s_)
            else:
                dpad = -(-d // 4) * 4  # the row-major FP32 rows, padded to 16 bytes
            paths = ["launched", "persistent"] if form != "fp32" else ["launched"]
            for path, nq in [(pa, nq_) for pa in paths for nq_ in nqs]:
                idx.set_persistent(max(nqs) if path == "persistent" else 0)
                q = ca.gen_uniform(nq, d, 9100 + nq, -1.0, 1.0)
                dq = ca.DeviceArray.from_host(q)
                dI = ca.DeviceArray((nq, k), np.int32)
                dD = ca.DeviceArray((nq, k), np.float32)
                for _ in range(3):
                    idx.search(dq, k, dI, dD)
                tot, full, rows, wall = [], [], 0, []
                for _ in range(args.sb_steps):
                    t0 = time.perf_counter()
                    idx.search(dq, k, dI, dD)
                    wall.append((time.perf_counter() - t0) * 1000.0)
                    st = idx.last_stats()
                    tot.append(st["ms_total"])
                    full.append(st["ms_filter_full"])
                    rows = int(st["rows_filter_full"])
                fms = float(np.median(full))
                bytes_full = rows * (dpad * elem + 4)
                gbs = bytes_full / (fms / 1000.0) / 1e9
                assert bool(st["persistent"]) == (path == "persistent")
                pt = {"precision": form, "path": path, "nq": nq, "k": k,
                      "ms_search": round(float(np.median(tot)), 4),
                      "ms_wall": round(float(np.median(wall)), 4), "ms_filter_full": round(fms, 4),
                      "rows_filter_full": rows, "bytes_filter_full": bytes_full,
                      "filter_gbs": round(gbs, 1), "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
                      "qps": round(nq / (float(np.median(tot)) / 1000.0), 1)}
                points.append(pt)
                log(f"[smallbatch] {n}x{d} {form} {path} nq={nq}: search {pt['ms_search']} ms, full filter "
                    f"{fms:.4f} ms = {gbs:.0f} GB/s ({gbs / HBM_PEAK_GBS:.3f} of peak)")
        print(json.dumps({"metric": "database stream GB/s of the full-level filter, small batches",
                          "unit": "GB/s", "n_gpus": 1, "higher_is_better": True,
                          "config": {"workload": f"{n}x{d} U[-1,1) index, k={k}", "n": n, "d": d},
                          "roofline": {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s"},
                          "points": points}), flush=True)
        idx.close()
    return 0


def mfma_peak(kind: str, random_operands: bool) -> float:
    """Measured MFMA throughput (TFLOP/s) from the benchmark-only probe library
    (curplsh_amd/bench/probe.hip): back-to-back MFMAs, no memory traffic."""
    import ctypes as C

    lib = C.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "curplsh_amd",
                              "libcurplsh_probe.so"))
    v = C.c_double(0.0)
    st = lib.rplsh_bench_mfma_peak(C.c_int({"bf16": 0, "fp8": 1, "fp32": 2}[kind]),
                                   C.c_int(1 if random_operands else 0), C.byref(v))
    if st != 0:
        raise RuntimeError(f"rplsh_bench_mfma_peak failed with status {st}")
    return float(v.value)


def copy_bandwidth(ca, nbytes: int, reps: int) -> float:
    """Read + write GB/s of a device-to-device copy of nbytes."""
    src = ca.DeviceArray((nbytes // 4,), np.float32)
    dst = ca.DeviceArray((nbytes // 4,), np.float32)
    ev0, ev1 = ca.Event(), ca.Event()
    ca.call("rplsh_memcpy_async", dst.ptr, src.ptr, nbytes, None)
    ca.call("rplsh_device_sync")
    ev0.record(0)
    for _ in range(reps):
        ca.call("rplsh_memcpy_async", dst.ptr, src.ptr, nbytes, None)
    ev1.record(0)
    ev1.synchronize()
    ms = ev0.elapsed_ms(ev1) / reps
    return 2.0 * nbytes / (ms / 1000.0) / 1e9


def copy_probe(args, ca) -> int:
    """hipMemcpy device-to-device of a known size: achieved HBM bandwidth
    (read + write bytes / time) and the reference dispatch whose FETCH_SIZE /
    WRITE_SIZE counters scale the filter's."""
    nbytes = int(args.copy_gib * (1 << 30))
    src = ca.DeviceArray((nbytes // 4,), np.float32)
    dst = ca.DeviceArray((nbytes // 4,), np.float32)
    ev0, ev1 = ca.Event(), ca.Event()
    ca.call("rplsh_memcpy_async", dst.ptr, src.ptr, nbytes, None)
    ca.call("rplsh_device_sync")
    reps = max(1, args.steps)
    ev0.record(0)
    for _ in range(reps):
        ca.call("rplsh_memcpy_async", dst.ptr, src.ptr, nbytes, None)
    ev1.record(0)
    ev1.synchronize()
    ms = ev0.elapsed_ms(ev1) / reps
    gbs = 2.0 * nbytes / (ms / 1000.0) / 1e9
    print(json.dumps({"metric": "device copy bandwidth", "value": round(gbs, 1), "unit": "GB/s",
                      "bytes_per_copy": nbytes, "ms_per_copy": round(ms, 4), "peak": HBM_PEAK_GBS,
                      "frac": round(gbs / HBM_PEAK_GBS, 4)}), flush=True)
    return 0


def serving(args, ca) -> int:
    """64 resident indexes, n log-uniform in [1e4, 2e6], d from {64,96,128,256},
    U[-1,1); requests with Poisson arrivals at --rate, each to a random index
    with nq log-uniform in [1, 2048] and k from {1, 10, 100}; continuous
    batching with HBM-budget admission (rplsh_server)."""
    import threading

    rng = np.random.default_rng(12345)
    res = ca.DeviceResources(0)
    idxs, dims = [], []
    t0 = time.time()
    total_rows = 0
    for i in range(args.indexes):
        n = int(np.exp(rng.uniform(np.log(1e4), np.log(2e6))))
        d = int(rng.choice([64, 96, 128, 256]))
        x = ca.gen_uniform(n, d, 70_000 + i, -1.0, 1.0)
        ix = ca.IndexBF(res, d, ca.IndexBFConfig(precision=args.serving_precision))
        ix.add(x)
        del x
        idxs.append(ix)
        dims.append(d)
        total_rows += n
    log(f"[serving] {args.indexes} indexes, {total_rows} rows, built in {time.time() - t0:.1f}s")
    reqs = []
    for r in range(args.requests):
        i = int(rng.integers(args.indexes))
        nq = int(np.clip(np.floor(np.exp(rng.uniform(0.0, np.log(2049.0)))), 1, 2048))
        k = int(rng.choice([1, 10, 100]))
        reqs.append((i, ca.gen_uniform(nq, dims[i], 900_000 + r, -1.0, 1.0), k))
    gaps = rng.exponential(1.0 / args.rate, size=args.requests)
    srv = ca.Server(slots=args.slots, persistent_max_queries=args.persistent_max_queries)
    ids = [srv.add_index(ix) for ix in idxs]
    # warm-up: one small request per index (workspaces, code paths)
    for i in range(args.indexes):
        srv.wait(srv.submit(ids[i], ca.gen_uniform(1, dims[i], 5, -1.0, 1.0), 10), timeout_s=120.0)
    srv.reset_stats()
    handles = [None] * len(reqs)

    def submitter():
        nxt = time.perf_counter()
        for j, (i, q, k) in enumerate(reqs):
            nxt += gaps[j]
            dt = nxt - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            handles[j] = srv.submit(ids[i], q, k)

    t0 = time.perf_counter()
    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    infos = []
    for h in handles:
        infos.append(srv.wait(h, timeout_s=600.0)[2])
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    nqs = sum(q.shape[0] for _, q, _ in reqs)
    log(f"[serving] stats {st}")
    line = {
        "metric": "serving QPS, 64 resident indexes, Poisson arrivals",
        "value": round(st["qps"], 1),
        "unit": "QPS",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": args.serving_precision,
        "data": "synthetic: seeded U[-1,1) indexes and queries",
        "config": {"workload": f"{args.indexes} indexes (n log-uniform 1e4..2e6, d in 64/96/128/256), "
                               f"{args.requests} requests, nq log-uniform 1..2048, k in 1/10/100",
                   "rate_rps": args.rate, "slots": args.slots, "rows": total_rows,
                   "persistent_max_queries": args.persistent_max_queries},
        "requests_per_s": round(len(reqs) / wall, 1),
        "queries": nqs,
        "p50_ms": round(st["p50_ms"], 3),
        "p99_ms": round(st["p99_ms"], 3),
        "mean_ms": round(st["mean_ms"], 3),
        "max_ms": round(st["max_ms"], 3),
        "batches": st["batches"],
        "device_ms_p50": round(float(np.percentile([inf["device_ms"] for inf in infos], 50)), 3),
        "device_ms_p99": round(float(np.percentile([inf["device_ms"] for inf in infos], 99)), 3),
        "queue_ms_p50": round(float(np.percentile([inf["queue_ms"] for inf in infos], 50)), 3),
        "graph_launches": st["graph_launches"],
        "persistent_runs": st["persistent_launches"],
        "persistent_batches": st["persistent_batches"],
        "graph_captures": st["graph_captures"],
        "co_batched_requests": sum(1 for inf in infos if inf["batch_requests"] > 1),
        "peak_inflight": st["peak_inflight"],
        "peak_reserved_gb": round(st["peak_reserved_bytes"] / 2**30, 3),
        "wall_s": round(wall, 3),
    }
    print(json.dumps(line), flush=True)
    return 0


# (tables, bits, probes, max_candidates, prerank) of the IndexRPLSH configurations
# timed; prerank f > 0 scores the candidates in BF16 and re-ranks only the f*k best
RPLSH_CONFIGS = [(8, 12, 0, 512, 0), (16, 14, 2, 2048, 0), (32, 12, 1, 4096, 0), (16, 20, 2, 1024, 0),
                 (32, 20, 4, 2048, 0), (32, 24, 8, 2048, 0), (64, 22, 4, 4096, 0),
                 (32, 20, 4, 2048, 8), (32, 24, 8, 2048, 8), (64, 22, 4, 4096, 8), (64, 22, 4, 4096, 16),
                 (64, 24, 4, 4096, 4), (64, 24, 4, 8192, 4), (64, 24, 8, 8192, 4), (128, 24, 4, 8192, 4)]


def rplsh_data(ca, name: str, n: int, nq: int):
    """(x, q) of an LSH workload, d = 128.
    c2: mixture of 256 centres U[-1,1)^128, sigma 0.1 (centre seed 2, point
        seeds 3 / 30): clusters of full intrinsic dimension;
    uniform16: U[-1,1)^16 (seeds 6 / 60) mapped to 128 dimensions by a fixed
        standard-normal 16x128 matrix / 4 (seed 7): unclustered data of
        intrinsic dimension 16, the regime hashing is meant for."""
    if name == "c2":
        return ca.gen_mixture(n, 128, 256, 0.1, 2, 3), ca.gen_mixture(nq, 128, 256, 0.1, 2, 30)
    a = np.random.default_rng(7).standard_normal((16, 128)).astype(np.float32) / np.float32(4.0)
    z = ca.gen_uniform(n, 16, 6, -1.0, 1.0)
    zq = ca.gen_uniform(nq, 16, 60, -1.0, 1.0)
    return np.ascontiguousarray(z @ a), np.ascontiguousarray(zq @ a)


def rplsh_front_end(args, ca) -> int:
    """IndexRPLSH (LSH candidates + exact re-rank) at 1M x 128, 10k queries,
    k = 10, on two workloads (rplsh_data). Recall against the plain path's
    exact result (bit-exact to the canonical oracle); the exact BF16 and FP32
    brute-force searches timed beside it."""
    d, nq, k = 128, args.nq, 10
    res = ca.DeviceResources(0)
    dq = ca.DeviceArray((nq, d), np.float32)
    dI = ca.DeviceArray((nq, k), np.int32)
    dD = ca.DeviceArray((nq, k), np.float32)

    def timed(index, stat_key):
        for _ in range(max(1, args.warmup)):
            index.search(dq, k, dI, dD)
        ms = []
        for _ in range(args.steps):
            index.search(dq, k, dI, dD)
            ms.append(index.last_stats()[stat_key])
        return float(np.median(ms)), dI.to_host()

    workloads = {}
    for name in args.lsh_data.split(","):
        n = 1_000_000 if name == "c2" else args.lsh_n
        x, q = rplsh_data(ca, name, n, nq)
        dq.copy_from(q)
        brute = {}
        bf = ca.IndexBF(res, d)
        bf.add(x)
        for form in ("fp32", "bf16"):
            bf.set_precision(form, 4)
            ms, I = timed(bf, "ms_total")
            brute[form] = {"ms": round(ms, 3), "qps": round(nq / ms * 1e3, 1)}
            if form == "fp32":
                truth = I
            else:
                brute[form]["recall_at_10"] = round(float(ca_recall(I, truth)), 5)
            log(f"[rplsh] {name} brute force {form}: {brute[form]}")
        bf.close()
        del bf
        rows = []
        configs = ([tuple(int(v) for v in c.split(":")) for c in args.lsh_configs.split(",")]
                   if args.lsh_configs else RPLSH_CONFIGS)
        for cfg in configs:
            L, K, P, B = cfg[:4]
            f = cfg[4] if len(cfg) > 4 else 0
            ix = ca.IndexRPLSH(res, d, ca.IndexRPLSHConfig(tables=L, bits=K, probes=P, max_candidates=B,
                                                           seed=1, prerank=f))
            t0 = time.perf_counter()
            ix.add(x)
            build_s = time.perf_counter() - t0
            ms, I = timed(ix, "ms_total")
            st = ix.last_stats()
            r = {"tables": L, "bits": K, "probes": P, "max_candidates": B, "prerank": f, "ms": round(ms, 3),
                 "qps": round(nq / ms * 1e3, 1), "recall_at_10": round(float(ca_recall(I, truth)), 5),
                 "raw_per_query": round(st["raw_candidates"] / nq, 1),
                 "distinct_per_query": round(st["distinct_candidates"] / nq, 1),
                 "ms_candidates": round(st["ms_candidates"], 3), "ms_rerank": round(st["ms_rerank"], 3),
                 "build_s": round(build_s, 3)}
            log(f"[rplsh] {name} {r}")
            rows.append(r)
            ix.close()
        workloads[name] = {"n": n, "brute_force": brute, "configs": rows}
        del x, q
    last = list(workloads)[-1]
    # the fastest config with recall@10 >= 0.9, else the highest-recall one
    best = max(workloads[last]["configs"],
               key=lambda r: (r["recall_at_10"] >= 0.9, r["qps"] if r["recall_at_10"] >= 0.9 else r["recall_at_10"]))
    line = {
        "metric": "IndexRPLSH QPS, k=10 (LSH candidates + exact re-rank), fastest config "
                  "with recall@10 >= 0.9 (else the highest-recall config)",
        "value": best["qps"],
        "unit": "QPS",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic: seeded Gaussian mixture (c2) and a 16-dimensional uniform subspace (uniform16)",
        "config": {"workload": f"{workloads[last]['n']}x128 {last}, {nq} queries, k=10", "best": best},
        "workloads": workloads,
    }
    print(json.dumps(line), flush=True)
    return 0


DATA_DESC = {
    "gaussian": "synthetic: seeded unit-normalised Gaussian vectors (counter-based generator)",
    "uniform01": "synthetic: seeded U[0,1) vectors (counter-based generator)",
    "mixture": "synthetic: seeded mixture of 256 centres U[-1,1) with N(0, 0.1^2) noise",
    "mixture1024": "synthetic: seeded mixture of 1024 centres U[-1,1) with N(0, 0.05^2) noise, "
                   "5% of rows exact copies of other rows",
}


def headline_data(ca, name: str, n: int, nq: int, d: int):
    """(x, q) of the headline workload: the metric's unit Gaussian vectors, or
    the uniform / clustered sets of the smaller configurations."""
    if name == "uniform01":
        return ca.gen_uniform(n, d, 0, 0.0, 1.0), ca.gen_uniform(nq, d, 1, 0.0, 1.0)
    if name == "mixture":
        return ca.gen_mixture(n, d, 256, 0.1, 2, 3), ca.gen_mixture(nq, d, 256, 0.1, 2, 30)
    if name == "mixture1024":
        return ca.gen_mixture(n, d, 1024, 0.05, 6, 60, 0.05), ca.gen_mixture(nq, d, 1024, 0.05, 6, 7)
    return ca.gen_gaussian(n, d, 4, True), ca.gen_gaussian(nq, d, 5, True)


def ca_recall(I, truth) -> float:
    """recall@k of I against truth (mean |I_i & T_i| / k; -1 never matches)."""
    k = truth.shape[1]
    hits = 0
    for a, b in zip(I, truth):
        hits += len(set(a[a >= 0].tolist()) & set(b[b >= 0].tolist()))
    return hits / (k * truth.shape[0])


if __name__ == "__main__":
    sys.exit(main())
