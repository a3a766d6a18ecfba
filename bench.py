"""Benchmark: Eclat on 1,000,000 transactions x 1,000 items (one MI355X).

A step is one pass of the support-counting path over the resident input:
the vertical store is built from the COO pairs already in HBM, then the
device search (level-1 counts, level-2 FP4 Gram on the matrix cores,
levels >= 3 on the support kernel with on-device candidate generation) runs
to completion and the pattern forest is copied back.  ``value`` is itemsets
per second of that step.  The JSON line also carries the support kernel's
roofline (algorithmic bitset bytes per launch / mean launch time, from HIP
events on the library's stream) and a CPU baseline: the oracle's OpenMP
vertical Eclat on the same workload, which doubles as the parity check.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                    [--workload eclat_1m_1k|pfeclat_2m_500] [--copy-calibration]

``--workload pfeclat_2m_500`` measures the periodic-frequent search on
2,000,000 timestamped baskets over 500 items (minSup 10,000, maxPer 2,000)
instead.  ``--copy-calibration`` adds one 256 MiB device-to-device copy of
known size, so HBM counters read under a profiler can be scaled to bytes.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "itemsets/s + bitset GB/s (Eclat 1M×1k), batched vs plain; serve p99 latency"
PF_METRIC = "itemsets/s + bitset GB/s (PF-Eclat 2M×500)"
HBM_PEAK_GBS = 8000.0
MINSUP = 1000  # 0.001 x 1,000,000
PF_MINSUP, PF_MAXPER = 10_000, 2_000  # 0.005 x 2,000,000; maxPer 2,000


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="eclat_1m_1k",
                    choices=["eclat_1m_1k", "eclat_4m_1k", "pfeclat_2m_500", "batch_256", "serve", "ingest_1m_1k", "roofs",
                             "output_1m_1k", "quickstart"])
    ap.add_argument("--requests", type=int, default=10_000, help="serve: number of requests")
    ap.add_argument("--qps", type=float, default=1000.0, help="serve: mean arrival rate")
    ap.add_argument("--pool", type=int, default=48, help="serve: databases the requests are drawn against")
    ap.add_argument("--slots", type=int, default=3, help="serve: concurrent searches (each its own stream)")
    ap.add_argument("--hbm-budget-gb", type=float, default=8.0, help="serve: admission budget")
    ap.add_argument("--p99-bound-ms", type=float, default=25.0, help="serve: stated p99 latency bound")
    ap.add_argument("--cpu-sample-s", type=float, default=15.0, help="serve CPU baseline: seconds of oracle work")
    ap.add_argument("--cpu-runs", type=int, default=5, help="CPU baseline: timed runs (median)")
    ap.add_argument("--no-extras", action="store_true",
                    help="eclat_1m_1k: skip the batched and serving summaries")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="launch level steps directly")
    ap.add_argument("--group", type=int, default=32, help="entries per support work item")
    ap.add_argument("--tile", type=int, default=1024, help="words per support tile (1024|2048)")
    ap.add_argument("--batched-only", action="store_true",
                    help="batch_256: the batched steps alone, on the schedule --group/--tile (no autotune, "
                         "no one-at-a-time or per-class comparison; for profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(split by prefix class after level 2, each GPU would mine its own classes from a replicated "
              "store).", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch_256":
        print(json.dumps(run_batch(args)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(run_serve(args)), flush=True)
        return
    if args.workload == "quickstart":
        line = run_quickstart(args)
        print(json.dumps(line), flush=True)
        if line["parity"] is False:
            sys.exit(1)
        return
    if args.workload == "output_1m_1k":
        line = run_output(args)
        print(json.dumps(line), flush=True)
        if line["parity"] is False:
            sys.exit(1)
        return
    if args.workload == "roofs":
        print(json.dumps(run_roofs(args)), flush=True)
        return
    if args.workload == "ingest_1m_1k":
        line = run_ingest(args)
        print(json.dumps(line), flush=True)
        if line["parity"] is False:
            sys.exit(1)
        return

    line, parity = run_search(args, args.workload)
    if args.workload == "eclat_1m_1k" and not args.no_extras:
        pf, pf_parity = run_search(argparse.Namespace(**{**vars(args), "steps": 10, "warmup": 2}), "pfeclat_2m_500")
        line["pfeclat_2m_500"] = {k: pf[k] for k in ("value", "unit", "ms_per_step", "config", "roofline",
                                                      "cpu_baseline", "parity")}
        line["batched"], line["serve"] = extras(args)
        line["quickstart"] = quickstart()
        if any(v["patterns"] != 1275 for k, v in line["quickstart"].items() if k != "workload"):
            parity = False
        if False in (line["batched"]["parity"], line["serve"]["parity"], pf_parity):
            parity = False
    print(json.dumps(line), flush=True)
    if parity is False:
        sys.exit(1)


# sources whose code a kernel's counter summary was measured on
SUPPORT_SOURCES = ["cupami_amd/csrc/support.hip", "cupami_amd/csrc/support_kernels.h", "cupami_amd/csrc/search.hip",
                   "cupami_amd/csrc/core.hip", "cupami_amd/csrc/cpm_internal.h"]
GRAM_SOURCES = ["cupami_amd/csrc/gram.hip", "cupami_amd/csrc/cpm_internal.h"]
MALL_BYTES = 256 << 20


def support_roofline(workload, kernel, stats, steps, store_bytes):
    """The support kernel against the level of the memory hierarchy that
    bounds it.  ``achieved``: algorithmic bytes per launch (summed in the
    kernel) / mean launch time (kernel clock stamps, live; launches past a
    search's last level exit at once and are not counted).  From the
    committed summaries: the kernel's L2 request bytes and fabric read bytes
    per search (counters) and the measured roofs.  Each level's time floor
    for a search is its bytes / its roof (L2 for requests; the Infinity Cache
    for fabric reads when the store fits it, HBM otherwise); the larger floor
    names the bound, and ``peak`` is that floor in algorithmic bytes, so
    ``frac`` = floor / kernel time per search <= 1 when the model holds.
    ``traffic``: HBM bytes per launch from FETCH_SIZE and WRITE_SIZE scaled by
    a copy of known size (an upper bound when the store sits in the Infinity
    Cache)."""
    from cupami_amd import profiling
    launches = max(stats["launches"], 1)
    avg_ms = stats["kernel_ms"] / launches
    alg = stats["bytes"] / launches
    achieved = alg / (avg_ms / 1e3) / 1e9 if avg_ms > 0 else 0.0
    line = {"bound": None, "achieved": round(achieved, 1), "peak": None, "unit": "GB/s", "frac": None,
            "traffic": None, "kernel": kernel, "avg_launch_ms": round(avg_ms, 4), "bytes_per_launch": alg,
            "launches_per_step": launches / steps}
    roofs = profiling.read_summary(ROOT / "profiles" / "roofs.md")
    cnt = profiling.read_summary(ROOT / "profiles" / f"{workload}_{kernel}_counters.md")
    if not roofs or "l2_request_bytes_per_search" not in cnt or avg_ms <= 0:
        line["basis"] = "no current counter summary for these sources: bound not placed"
        return line
    l2_b = float(cnt["l2_request_bytes_per_search"])
    fab_b = float(cnt["fabric_read_bytes_per_search"])
    fab_level = "mall" if store_bytes <= MALL_BYTES else "hbm"
    r_l2, r_fab = float(roofs["l2_gb_s"]), float(roofs[f"{fab_level}_gb_s"])
    t_l2, t_fab = l2_b / (r_l2 * 1e9), fab_b / (r_fab * 1e9)
    bound, floor = ("l2", t_l2) if t_l2 >= t_fab else (fab_level, t_fab)
    t = stats["kernel_ms"] / steps / 1e3    # kernel time per search
    alg_s = stats["bytes"] / steps
    peak = alg_s / floor / 1e9
    per_launch = launches / steps
    line.update(bound=bound, peak=round(peak, 1), frac=round(achieved / peak, 4),
                traffic=(float(cnt["hbm_bytes_per_search"]) / per_launch
                         if "hbm_bytes_per_search" in cnt else None))
    line["basis"] = {
        "l2_request_bytes_per_search": l2_b, "l2_hit_rate": float(cnt.get("l2_hit_rate", "nan")),
        "l2_roof_gb_s": r_l2, "l2_gb_s": round(l2_b / t / 1e9, 1),
        "fabric_read_bytes_per_search": fab_b, f"{fab_level}_roof_gb_s": r_fab,
        "fabric_gb_s": round(fab_b / t / 1e9, 1),
        "fabric_frac_of_hbm_published": round(fab_b / t / 1e9 / HBM_PEAK_GBS, 4),
        "kernel_ms_per_search": round(t * 1e3, 4),
        "summaries": ["profiles/roofs.md", f"profiles/{workload}_{kernel}_counters.md"]}
    return line


INGEST_SOURCES = ["cupami_amd/csrc/ingest.hip", "cupami_amd/csrc/core.hip", "cupami_amd/csrc/cpm_internal.h"]


def ingest_traffic():
    """HBM bytes per launch of k_ingest_tile from its committed counter
    summary (None once the sources it measured change)."""
    from cupami_amd import profiling
    cnt = profiling.read_summary(ROOT / "profiles" / "ingest_1m_1k_k_ingest_tile_counters.md")
    return float(cnt["hbm_bytes_per_launch"]) if "hbm_bytes_per_launch" in cnt else None


def cpu_model():
    """The host CPU's model name, logical CPU count and the oracle's threads."""
    import oracle
    model = ""
    try:
        model = next((l.split(":", 1)[1].strip() for l in Path("/proc/cpuinfo").read_text().splitlines()
                      if l.startswith("model name")), "")
    except OSError:
        pass
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "threads": oracle.threads()}


def timed_median(fn, runs):
    """(median seconds, every run's seconds, the last result) of ``runs``
    calls of ``fn``."""
    times, out = [], None
    for _ in range(runs):
        t = time.perf_counter()
        out = fn()
        times.append(time.perf_counter() - t)
    return float(sorted(times)[len(times) // 2]), [round(x, 3) for x in times], out


def run_search(args, workload):
    """One database's store build + device search per step (eclat_1m_1k,
    eclat_4m_1k, pfeclat_2m_500), its roofline and CPU baseline."""
    import numpy as np
    import torch
    from cupami_amd import _hip, datagen

    periodic = workload == "pfeclat_2m_500"
    big = workload == "eclat_4m_1k"
    t = time.perf_counter()
    db = (datagen.temporal_2m_500() if periodic else
          datagen.transactional(4_000_000, 1000, 10.0, 1.1, 202) if big else datagen.eclat_1m_1k())
    labels, cols = datagen.label_columns(db.items)
    minsup = PF_MINSUP if periodic else 4 * MINSUP if big else MINSUP
    end = int(db.timestamps[-1]) if periodic else 0
    n, items, nnz = db.n_transactions, len(labels), db.rows.size
    log(f"bench: {workload}: generated {n} x {items} ({nnz} pairs) in {time.perf_counter() - t:.1f}s")

    dev = torch.device("cuda:0")
    d_rows = torch.from_numpy(db.rows.view(np.int64)).to(dev)
    d_cols = torch.from_numpy(cols.view(np.int32)).to(dev)
    d_ts = torch.from_numpy(db.timestamps.astype(np.int64)).to(dev) if periodic else None
    torch.cuda.synchronize()
    ctx = _hip.context(0)
    ctx.set_search_options(args.group, args.tile, not args.no_graph, True)

    def step():
        bits = ctx.build_bits_device(d_rows.data_ptr(), d_cols.data_ptr(), nnz, n, items)
        try:
            if periodic:
                return ctx.pfeclat_mine_device(bits, d_ts.data_ptr(), end, minsup, PF_MAXPER)
            return ctx.eclat_mine(bits, minsup, use_gram=True)
        finally:
            bits.close()

    for i in range(args.warmup):
        pats = step()
        log(f"bench: warmup {i} -> {len(pats)} patterns")

    ctx.reset_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pats = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    stats = ctx.stats()
    n_pat = len(pats)
    ms_per_step = elapsed / args.steps * 1e3
    value = n_pat / (elapsed / args.steps)
    log(f"bench: {args.steps} steps, {ms_per_step:.2f} ms/step, {n_pat} patterns")

    # the reference's mine() also builds its pattern dictionary: the same
    # step followed by getPatterns()'s dict, timed apart from the headline
    def step_with_dict():
        p = step()
        return dict(zip(p.keys(labels), p.support.tolist()))
    dict_s, _, got_dict = timed_median(step_with_dict, 3)

    kernel = "k_support_periodic" if periodic else "k_support_count"
    roofline = support_roofline(workload, kernel, stats, args.steps, items * (((n + 63) // 64 + 63) // 64 * 64) * 8)
    gram_ms = stats["gram_ms"] / max(stats["gram_launches"], 1)
    gram_ops = stats["gram_ops"] / max(stats["gram_launches"], 1)
    # dense FP4 MFMA peak of the MI355X (the Gram's operands are E2M1 0 / 1)
    gram = {"kernel": "k_gram", "dtype": "fp4", "avg_launch_ms": round(gram_ms, 4),
            "tops": round(gram_ops / (gram_ms / 1e3) / 1e12, 2) if gram_ms > 0 else 0.0,
            "peak_tops": 10100.0}
    gram["frac"] = round(gram["tops"] / gram["peak_tops"], 4)
    from cupami_amd import profiling
    roofs = profiling.read_summary(ROOT / "profiles" / "roofs.md")
    if roofs.get("mfma_fp4_32x32x64_tops"):
        gram["measured_loop_tops"] = float(roofs["mfma_fp4_32x32x64_tops"])
        gram["frac_of_loop"] = round(gram["tops"] / gram["measured_loop_tops"], 4)
    gcnt = profiling.read_summary(ROOT / "profiles" / f"{workload}_k_gram_counters.md")
    if gcnt.get("avg_sq_valu_mfma_busy_cycles"):
        # busy cycles summed over every SIMD; GRBM_GUI_ACTIVE summed over the 8 XCDs
        simds = 4 * int(roofs.get("cu_count", 256) or 256)
        gram["mfma_busy_frac"] = round(float(gcnt["avg_sq_valu_mfma_busy_cycles"]) /
                                       (float(gcnt["avg_grbm_gui_active"]) / 8 * simds), 4)

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        import oracle
        host = oracle.bits_build(db.rows, cols, n, items)
        runs = args.cpu_runs
        if periodic:
            search = lambda: oracle.pfeclat(host, db.timestamps, n, end, minsup, PF_MAXPER)
            expect = search()   # parity first, then the timed runs
            parity = pats.to_dict() == expect
            med, all_s, _ = timed_median(search, runs)
            cpu = {"value": round(len(expect) / med, 1), "unit": "itemsets/s", "kind": "port",
                   "sample": f"whole workload ({len(expect)} itemsets), median of {runs} warm runs after the "
                             "parity check: oracle/cpm_oracle.c vertical PF-Eclat, OpenMP over prefix classes",
                   "median_s": round(med, 3), "runs_s": all_s}
        else:
            eclat = lambda: oracle.eclat(host, minsup)
            apriori = lambda: oracle.apriori(host, minsup)
            expect = eclat()
            parity = pats.to_dict() == expect and oracle.apriori(host, minsup) == expect
            if not big:
                labelled = {"\t".join(labels[c] for c in k): v for k, v in expect.items()}
                parity = parity and got_dict == labelled
            e_med, e_all, _ = timed_median(eclat, runs)
            a_med, a_all, _ = timed_median(apriori, runs)
            best = min(e_med, a_med)
            cpu = {"value": round(len(expect) / best, 1), "unit": "itemsets/s", "kind": "port",
                   "sample": f"whole workload ({len(expect)} itemsets), median of {runs} warm runs each after the "
                             "parity check; value from the faster of the clApriori restatement "
                             "(u32 words, batches of 4096, _opencl_bitsets.py:21-42, clApriori.py:164-196) and "
                             "the vertical Eclat oracle, both oracle/cpm_oracle.c with OpenMP",
                   "clapriori_restatement": {"median_s": round(a_med, 3), "runs_s": a_all,
                                             "itemsets_per_s": round(len(expect) / a_med, 1)},
                   "eclat_oracle": {"median_s": round(e_med, 3), "runs_s": e_all,
                                    "itemsets_per_s": round(len(expect) / e_med, 1)}}
        cpu.update(cpu_model())
        cpu["cores"] = cpu["threads"]
        log(f"bench: cpu baseline {cpu['value']} itemsets/s on {cpu['threads']} threads, parity={parity}")
        if not parity:
            log("bench: PARITY FAILURE against the oracle")

    if args.copy_calibration:
        src = torch.empty(256 << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("bench: copied 268435456 bytes device to device (counter calibration)")

    if periodic:
        config = {"workload": "pfeclat_2m_500", "transactions": n, "items": items, "lambda": 8,
                  "zipf_s": 1.05, "minsup": minsup, "maxper": PF_MAXPER, "patterns": n_pat,
                  "parallelism": "single GPU"}
    else:
        config = {"workload": workload, "transactions": n, "items": items,
                  "lambda": 10, "zipf_s": 1.1, "minsup": minsup, "patterns": n_pat,
                  "parallelism": "single GPU"}
    line = {
        "metric": PF_METRIC if periodic else METRIC, "value": round(value, 1), "unit": "itemsets/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": f"synthetic (numpy PCG64 seed {404 if periodic else 202})",
        "config": config,
        "roofline": roofline, "cpu_baseline": cpu, "gram": gram,
        "with_get_patterns": {"ms_per_step": round(dict_s * 1e3, 3), "itemsets_per_s": round(n_pat / dict_s, 1),
                              "note": "the step followed by getPatterns()'s dict, median of 3"},
        "levels_per_step": stats["levels"], "parity": parity,
    }
    return line, parity


def quickstart():
    """The reference's own recorded case: its quickstart database
    (TransactionalDatabase(10000, 50, avg_length=6, seed=42), restated in
    tests/golden/generator.py) mined at minSup 0.01 -- 1275 patterns for
    every miner.  getMiningTime() here beside the notebook's recorded times
    (quickstart.ipynb:86-88,127-129,168-170: one first call each, NVRTC JIT
    included, unstated GPU), and mine() with the _finalPatterns dict built,
    the scope of the reference's mining time."""
    sys.path.insert(0, str(ROOT / "tests" / "golden"))
    import generator
    from cupami_amd import cuApriori, cuEclat, cuFPGrowth
    rows = generator.transactional(10000, 50, 6, 42)
    recorded = {"cuApriori": 718.0, "cuEclat": 254.5, "cuFPGrowth": 608.9}
    out = {"workload": "quickstart 10000x50, minSup 0.01 (the reference's generator, seed 42)"}
    for cls in (cuApriori, cuEclat, cuFPGrowth):
        times, full, n = [], [], None
        for _ in range(5):
            m = cls(rows, 0.01)
            t = time.perf_counter()
            m.mine()
            # the reference's mine() ends with its _finalPatterns dict built
            # (label tuple -> support); here that dict is built on first use
            n = len(m._finalPatterns)
            full.append((time.perf_counter() - t) * 1e3)
            times.append(m.getMiningTime() * 1e3)
        out[cls.__name__] = {"patterns": n, "first_ms": round(times[0], 2), "best_ms": round(min(times), 2),
                             "with_final_patterns_best_ms": round(min(full), 2),
                             "reference_notebook_ms": recorded[cls.__name__]}
    return out


def run_quickstart(args):
    """The reference's recorded case on its own (see quickstart()), with the
    three miners' dictionaries checked against each other and against the
    per-basket subset count of the reference's tests
    (test_transaction_formats.py:15-23), which is also the CPU baseline."""
    q = quickstart()
    sys.path.insert(0, str(ROOT / "tests" / "golden"))
    import generator
    from cupami_amd import cuApriori, cuEclat, cuFPGrowth
    rows = generator.transactional(10000, 50, 6, 42)
    got = []
    for cls in (cuApriori, cuEclat, cuFPGrowth):
        m = cls(rows, 0.01)
        m.mine()
        got.append(m.getPatterns())
    parity = got[0] == got[1] == got[2] and len(got[0]) == 1275
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        t = time.perf_counter()
        expect = oracle.brute_force_rows(rows, oracle.support_threshold(0.01, len(rows)))
        cpu_s = time.perf_counter() - t
        parity = parity and got[1] == expect
        cpu = {"value": round(len(expect) / cpu_s, 1), "unit": "itemsets/s", "cores": 1, "kind": "port",
               "sample": f"whole workload, one run ({cpu_s:.2f}s): the per-basket subset count of "
                         "test_transaction_formats.py:15-23 (pure Python)"}
    best = q["cuEclat"]["best_ms"]
    return {
        "metric": "getMiningTime() on the reference's quickstart case (itemsets/s, cuEclat)",
        "value": round(1275 / (best / 1e3), 1), "unit": "itemsets/s", "n_gpus": 1, "steps": 5, "warmup": 0,
        "ms_per_step": best, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "the reference's generator restated (seed 42)",
        "config": {"workload": "quickstart", "transactions": 10000, "items": 50, "avg_length": 6,
                   "minsup": 0.01, "parallelism": "single GPU"},
        "roofline": None, "cpu_baseline": cpu, "quickstart": q, "parity": parity,
    }


def extras(args):
    """Compact summaries of the batched (256 databases) and serving
    workloads for the headline line (their own workloads report in full);
    both check their results against the oracle."""
    sub = argparse.Namespace(**vars(args))
    sub.steps, sub.warmup, sub.no_cpu_baseline, sub.cpu_sample_s = 3, 1, False, 5.0
    b = run_batch(sub)
    batched = {"workload": "batch_256", "databases": b["config"]["databases"],
               "patterns": b["config"]["patterns"], "itemsets_per_s": b["value"],
               "ms_per_step": b["ms_per_step"], "one_at_a_time_ms": b["plain"]["one_at_a_time_ms"],
               "speedup_vs_one_at_a_time": b["plain"]["speedup_batched"],
               "speedup_vs_per_class_subset": b["plain"]["per_class_subset"]["speedup_batched"],
               "schedule": b["config"]["schedule"], "cpu_baseline": b["cpu_baseline"], "parity": b["parity"]}
    v = run_serve(sub)
    serve = {"workload": "serve", "requests": v["config"]["requests"], "qps": v["config"]["qps"],
             "pool_databases": v["config"]["pool_databases"], "slots": v["config"]["slots"],
             "p50_ms": v["served"]["p50_ms"], "p99_ms": v["served"]["p99_ms"],
             "p99_bound_ms": v["p99_bound_ms"], "p99_within_bound": v["p99_within_bound"],
             "itemsets_per_s": v["served"]["itemsets_per_s"],
             "one_at_a_time_p50_ms": v["one_at_a_time"]["p50_ms"],
             "one_at_a_time_p99_ms": v["one_at_a_time"]["p99_ms"],
             "one_at_a_time_itemsets_per_s": v["one_at_a_time"]["itemsets_per_s"],
             "one_at_a_time_resident_p99_ms": v["one_at_a_time_resident"]["p99_ms"],
             "batches": v["served"]["batches"], "max_batch": v["served"]["max_batch"],
             "slots_busy_max": v["served"]["slots_busy_max"], "cpu_baseline": v["cpu_baseline"],
             "parity": v["parity"]}
    return batched, serve


def per_class_plain(ctx, labels, bits, n, minsup):
    """The reference's schedule for one database on this GPU: level 1 and the
    pair supports counted in batches, then one launch per prefix class
    (cuEclat.py:168-245), driven from the host as cuEclat's host path does."""
    from cupami_amd.frequentPattern.cuEclat import cuEclat
    m = cuEclat.__new__(cuEclat)
    m._minSup = minsup
    m._transactionCount = n
    m._finalPatterns = {}
    m._mineHost(ctx, list(labels), bits)
    return m._finalPatterns


def run_batch(args):
    """256 databases mined together (one store build call, one search)
    against the same engine one database at a time, and against the
    reference's per-class launch schedule on a subset."""
    import numpy as np
    import torch
    from cupami_amd import _hip, datagen
    from cupami_amd.batch import BatchMiner

    t = time.perf_counter()
    dbs = datagen.batch_256()
    cols_all, labels_all = [], []
    for db, _ in dbs:
        labels, cols = datagen.label_columns(db.items)
        labels_all.append(labels)
        cols_all.append(cols)
    minsups = [ms for _, ms in dbs]
    nnz = [db.rows.size for db, _ in dbs]
    log(f"bench: generated {len(dbs)} databases ({sum(nnz)} pairs) in {time.perf_counter() - t:.1f}s")

    dev = torch.device("cuda:0")
    d_rows = torch.from_numpy(np.concatenate([db.rows for db, _ in dbs]).view(np.int64)).to(dev)
    d_cols = torch.from_numpy(np.concatenate(cols_all).view(np.int32)).to(dev)
    torch.cuda.synchronize()
    base_r, base_c = d_rows.data_ptr(), d_cols.data_ptr()
    specs, off = [], 0
    for (db, _), labels, k in zip(dbs, labels_all, nnz):
        specs.append((base_r + 8 * off, base_c + 4 * off, k, db.n_transactions, len(labels)))
        off += k

    ctx = _hip.context(0)
    miner = BatchMiner(ctx)
    if args.batched_only:
        config, table = (args.group, args.tile), {}
        miner.use(config)
    else:
        config, table = miner.tune(specs, minsups)
    log(f"bench: autotune {({f'{g}x{w}': round(v * 1e3, 2) for (g, w), v in table.items()})} -> {config}")
    # each step releases the previous result first, as a consumer that hands
    # results on would (its host blocks are then reused)
    pats = None
    for _ in range(args.warmup):
        pats = None
        pats = miner.mine(specs, minsups)
    ctx.reset_stats()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pats = None
        pats = miner.mine(specs, minsups)
    ctx.synchronize()
    batched_s = (time.perf_counter() - t0) / args.steps
    stats = ctx.stats()
    n_pat = len(pats)
    log(f"bench: batched {batched_s * 1e3:.2f} ms/step, {n_pat} patterns")
    if args.batched_only:
        return {"metric": "itemsets/s (batched Eclat, 256 databases)", "value": round(n_pat / batched_s, 1),
                "unit": "itemsets/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                "ms_per_step": round(batched_s * 1e3, 3), "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "u64", "data": "synthetic (numpy PCG64 seeds 303, 303000+i)",
                "config": {"workload": "batch_256", "databases": len(dbs), "patterns": n_pat,
                           "schedule": {"group": config[0], "tile_words": config[1]},
                           "parallelism": "single GPU"},
                "roofline": None, "cpu_baseline": None, "parity": None}

    # plain: the same engine, one database at a time
    def sequential(idx):
        out = []
        for i in idx:
            r, c, k, n, items = specs[i]
            bits = ctx.build_bits_device(r, c, k, n, items)
            try:
                out.append(ctx.eclat_mine(bits, minsups[i]))
            finally:
                bits.close()
        return out
    sequential(range(len(dbs)))
    ctx.synchronize()
    t = time.perf_counter()
    seq = sequential(range(len(dbs)))
    ctx.synchronize()
    seq_s = time.perf_counter() - t
    log(f"bench: one database at a time {seq_s * 1e3:.1f} ms")

    # the reference's per-class schedule, on every 16th database
    subset = list(range(0, len(dbs), 16))
    sub_specs = [specs[i] for i in subset]
    sub_ms = [minsups[i] for i in subset]
    miner.mine(sub_specs, sub_ms)
    ctx.synchronize()
    t = time.perf_counter()
    sub_batched = miner.mine(sub_specs, sub_ms)
    ctx.synchronize()
    sub_batched_s = time.perf_counter() - t
    t = time.perf_counter()
    per_class = []
    for i in subset:
        r, c, k, n, items = specs[i]
        bits = ctx.build_bits_device(r, c, k, n, items)
        try:
            per_class.append(per_class_plain(ctx, range(items), bits, n, minsups[i]))
        finally:
            bits.close()
    ctx.synchronize()
    per_class_s = time.perf_counter() - t
    log(f"bench: subset of {len(subset)}: batched {sub_batched_s * 1e3:.1f} ms, per-class {per_class_s * 1e3:.1f} ms")

    # parity: batched == one at a time (all), == per-class (subset), == oracle
    def same(tag, idx, got, want):
        if got == want:
            return True
        missing = [k for k in want if k not in got]
        extra = [k for k in got if k not in want]
        wrong = [k for k in want if k in got and got[k] != want[k]]
        log(f"bench: MISMATCH {tag} database {idx}: {len(got)} vs {len(want)} itemsets, "
            f"missing {len(missing)} {missing[:3]}, extra {len(extra)} {extra[:3]}, "
            f"support differs {len(wrong)} {[(k, got[k], want[k]) for k in wrong[:3]]}")
        return False

    parts = pats.split(len(dbs))
    parity = all([same("batched/one-at-a-time", i, p, s.to_dict()) for i, (p, s) in enumerate(zip(parts, seq))])
    sub_parts = sub_batched.split(len(subset))
    parity = all([same("batched-subset/per-class", subset[j], sub_parts[j],
                       {tuple(k): v for k, v in per_class[j].items()}) for j in range(len(subset))]) and parity
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        threads = oracle.threads()
        hosts = [oracle.bits_build(db.rows, c, db.n_transactions, len(l))
                 for (db, _), c, l in zip(dbs, cols_all, labels_all)]
        t = time.perf_counter()
        expect = [oracle.eclat(h, ms) for h, ms in zip(hosts, minsups)]
        cpu_s = time.perf_counter() - t
        parity = all([same("batched/oracle", i, p, e) for i, (p, e) in enumerate(zip(parts, expect))]) and parity
        n_exp = sum(len(e) for e in expect)
        cpu = {"value": round(n_exp / cpu_s, 1), "unit": "itemsets/s", "cores": threads, "kind": "port",
               "sample": f"all {len(dbs)} databases, one run ({n_exp} itemsets in {cpu_s:.2f}s): "
                         "oracle/cpm_oracle.c vertical Eclat, OpenMP over prefix classes"}
    log(f"bench: parity={parity}")

    store_bytes = sum(len(l) * (((db.n_transactions + 63) // 64 + 63) // 64 * 64) * 8
                      for (db, _), l in zip(dbs, labels_all))
    roofline = support_roofline("batch_256", "k_support_count", stats, args.steps, store_bytes)
    line = {
        "metric": "itemsets/s (batched Eclat, 256 databases) vs plain", "value": round(n_pat / batched_s, 1),
        "unit": "itemsets/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(batched_s * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic (numpy PCG64 seeds 303, 303000+i)",
        "config": {"workload": "batch_256", "databases": len(dbs), "transactions": sum(db.n_transactions for db, _ in dbs),
                   "pairs": sum(nnz), "patterns": n_pat, "schedule": {"group": config[0], "tile_words": config[1]},
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain": {"one_at_a_time_ms": round(seq_s * 1e3, 2),
                  "speedup_batched": round(seq_s / batched_s, 2),
                  "per_class_subset": {"databases": subset, "per_class_ms": round(per_class_s * 1e3, 1),
                                       "batched_ms": round(sub_batched_s * 1e3, 2),
                                       "speedup_batched": round(per_class_s / sub_batched_s, 1)}},
        "autotune": {f"{g}x{w}": round(v * 1e3, 3) for (g, w), v in table.items()},
        "parity": parity,
    }
    return line


# (name, buffer bytes, passes, XCD-sliced): AND + popcount of the buffer's two
# halves; sliced buffers give each XCD its own eighth, as the support kernel's
# tile == XCD mapping does
ROOF_CASES = (
    ("l2_2mib_per_xcd", 16 << 20, 20000, True),
    ("l2_4mib_per_xcd", 32 << 20, 10000, True),
    ("mall_16mib_per_xcd", 128 << 20, 1000, True),
    ("mall_25mib_per_xcd", 200 << 20, 600, True),
    ("hbm_4gib", 4 << 30, 8, True),
    ("hbm_8gib_flat", 8 << 30, 4, False),
)


def run_roofs(args):
    """This device's measured roofs: AND + popcount streaming rates with the
    data resident in L2, in the Infinity Cache and in HBM, and the FP4 MFMA
    loop rate (best of three timed launches after one warm launch each)."""
    import torch
    from cupami_amd import _hip
    ctx = _hip.context(0)
    dev = torch.device("cuda:0")
    sink = torch.zeros(1, dtype=torch.int64, device=dev)
    out = {}
    for name, nbytes, reps, sliced in ROOF_CASES:
        buf = torch.randint(0, 2**62, (nbytes // 8,), dtype=torch.int64, device=dev)
        torch.cuda.synchronize()
        ctx.roof_stream(buf.data_ptr(), nbytes, max(1, reps // 10), sliced, sink.data_ptr())
        best = None
        for _ in range(3):
            ms, nb = ctx.roof_stream(buf.data_ptr(), nbytes, reps, sliced, sink.data_ptr())
            gbs = nb / (ms / 1e3) / 1e9
            best = gbs if best is None else max(best, gbs)
        out[name] = {"buffer_bytes": nbytes, "passes": reps, "xcd_sliced": sliced, "gb_s": round(best, 1)}
        log(f"bench: roof {name}: {best:.1f} GB/s")
        del buf
    fs = torch.zeros(ctx_grid(ctx), dtype=torch.float32, device=dev)
    for shape, name in ((16, "mfma_fp4_16x16x128"), (32, "mfma_fp4_32x32x64")):
        ctx.roof_mfma(2000, fs.data_ptr(), shape)
        best = 0.0
        for _ in range(3):
            ms, ops = ctx.roof_mfma(20000, fs.data_ptr(), shape)
            best = max(best, ops / (ms / 1e3) / 1e12)
        out[name] = {"tops": round(best, 1), "published_dense_tops": 10100.0}
        log(f"bench: roof {name}: {best:.1f} TOP/s")
    # a phase boundary inside one launch (grid-wide barrier with an
    # agent-scope release) against a launch boundary in a graph
    for blocks in (256, 1024, 2048):
        ctx.roof_phases(blocks, 8, 4096)
        runs = [ctx.roof_phases(blocks, 64, 4096) for _ in range(3)]
        ub, ul = min(r[0] for r in runs), min(r[1] for r in runs)
        out[f"phase_{blocks}_blocks"] = {"us_per_barrier": round(ub, 2), "us_per_launch": round(ul, 2),
                                         "dirty_bytes_per_block": 4096}
        log(f"bench: roof phases, {blocks} workgroups: barrier {ub:.2f} us, graph launch {ul:.2f} us")
    props = torch.cuda.get_device_properties(0)
    return {"metric": "measured roofs (AND+popcount GB/s by residency; FP4 MFMA TOP/s)",
            "value": out["mall_16mib_per_xcd"]["gb_s"], "unit": "GB/s", "n_gpus": 1, "steps": 3, "warmup": 1,
            "higher_is_better": True, "device": props.name, "cu_count": props.multi_processor_count,
            "roofs": out, "hbm_published_gb_s": HBM_PEAK_GBS}


def ctx_grid(ctx):
    import torch
    return torch.cuda.get_device_properties(0).multi_processor_count * 8


def run_ingest(args):
    """The step before the search: the store built from a CUPAMI01 file's
    arrays (offsets u64[N+1], ids u32[M]) already in HBM, labels included.
    Baseline: the reference's host work for the same arrays (the BIN branch of
    _getDevicePairs, restated in the oracle) plus the oracle's store build."""
    import numpy as np
    import torch
    from cupami_amd import _hip, datagen

    t = time.perf_counter()
    db = datagen.eclat_1m_1k()
    n, m, universe = db.n_transactions, db.items.size, db.n_items
    offsets = np.searchsorted(db.rows, np.arange(n + 1, dtype=np.uint64)).astype(np.uint64)
    log(f"bench: generated {n} baskets, {m} ids in {time.perf_counter() - t:.1f}s")
    dev = torch.device("cuda:0")
    d_off = torch.from_numpy(offsets.view(np.int64)).to(dev)
    d_ids = torch.from_numpy(db.items.view(np.int32)).to(dev)
    torch.cuda.synchronize()
    ctx = _hip.context(0)

    def step():
        bits, labels = ctx.build_bits_csr_device(d_off.data_ptr(), n, d_ids.data_ptr(), m, universe)
        return bits, labels

    for _ in range(args.warmup):
        bits, labels = step()
        bits.close()
    ctx.reset_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bits, labels = step()
        if _ + 1 < args.steps:
            bits.close()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    st = ctx.ingest_stats()
    ms_per_step = elapsed / args.steps * 1e3
    _, n_items, _, stride = bits.shape()
    file_bytes = 8 * (n + 1) + 4 * m
    store_bytes = 8 * n_items * stride
    builds = max(st["builds"], 1)
    tile_ms = st["tile_ms"] / builds
    tile_bytes = st["tile_bytes"] / builds
    achieved = tile_bytes / (tile_ms / 1e3) / 1e9 if tile_ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": ingest_traffic(),
                "kernel": "k_ingest_tile",
                "avg_launch_ms": round(tile_ms, 4), "bytes_per_launch": tile_bytes}
    log(f"bench: {args.steps} builds, {ms_per_step:.3f} ms/step (device span {st['build_ms'] / builds:.3f} ms), "
        f"tile {tile_ms:.4f} ms at {achieved:.0f} GB/s")

    parity = None
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        t = time.perf_counter()
        labs, rr, cc, n_rows = oracle.csr_to_coo(offsets, db.items)
        host = oracle.bits_build(rr, cc, n_rows, len(labs))
        cpu_s = time.perf_counter() - t
        parity = ([str(v) for v in labels.tolist()] == labs
                  and np.array_equal(bits.download(), host))
        cpu = {"value": round(m / cpu_s, 1), "unit": "entries/s", "cores": 1, "kind": "port",
               "sample": f"whole workload, one run ({cpu_s:.2f}s): the BIN branch of _getDevicePairs "
                         "(numpy, oracle.csr_to_coo) then oracle/cpm_oracle.c's store build"}
        log(f"bench: cpu baseline {cpu_s:.2f}s, parity={parity}")
    bits.close()
    if args.copy_calibration:
        src = torch.empty(256 << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("bench: copied 268435456 bytes device to device (counter calibration)")

    # the same build with the arrays in host memory (PCIe copies included),
    # and the pairs path it replaces: host COO from the arrays, then the pairs build
    t = time.perf_counter()
    for _ in range(3):
        b, _l = ctx.build_bits_csr(offsets, db.items, universe)
        b.close()
    host_arrays_ms = (time.perf_counter() - t) / 3 * 1e3
    from cupami_amd.frequentPattern.cuApriori import binary_pairs
    t = time.perf_counter()
    labs, rr, cc = binary_pairs(offsets, db.items)
    b = ctx.build_bits(rr, cc, int(np.count_nonzero(np.diff(offsets))), len(labs))
    ctx.synchronize()
    pairs_ms = (time.perf_counter() - t) * 1e3
    b.close()
    return {
        "metric": "CUPAMI01 arrays -> bitset store (entries/s)", "value": round(m / (elapsed / args.steps), 1),
        "unit": "entries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic (numpy PCG64 seed 202)",
        "config": {"workload": "ingest_1m_1k", "baskets": n, "ids": m, "universe": universe,
                   "items": n_items, "file_bytes": file_bytes, "store_bytes": store_bytes,
                   "parallelism": "single GPU"},
        "roofline": roofline, "cpu_baseline": cpu,
        "build_device_ms": round(st["build_ms"] / builds, 4),
        "host_arrays_ms": round(host_arrays_ms, 3),
        "pairs_path_ms": round(pairs_ms, 1),
        "parity": parity,
    }


def run_output(args):
    """The step after the search: the 1M x 1k Eclat result (43,645 patterns)
    turned into getPatterns()'s dict and save()'s file by the library
    (cpm_patterns_keys / cpm_patterns_save), against the reference's Python
    rendering of the same result (cuApriori.py:537-552: a tuple per pattern,
    "\t".join per key, sorted() for the file)."""
    import tempfile
    import numpy as np
    import torch
    from cupami_amd import _hip, datagen

    db = datagen.eclat_1m_1k()
    labels, cols = datagen.label_columns(db.items)
    ctx = _hip.context(0)
    bits = ctx.build_bits(db.rows, cols, db.n_transactions, len(labels))
    pats = ctx.eclat_mine(bits, MINSUP, use_gram=True)
    bits.close()
    torch.cuda.synchronize()
    n = len(pats)
    tmp = Path(tempfile.mkdtemp())

    def native():
        d = dict(zip(pats.keys(labels), pats.support.tolist()))
        pats.save(tmp / "native.txt", labels)
        return d

    def python():
        tuples = pats.to_dict(labels)
        d = {"\t".join(p): s for p, s in tuples.items()}
        with (tmp / "python.txt").open("w", encoding="utf-8") as handle:
            for pattern, support in sorted(tuples.items()):
                handle.write(f"{chr(9).join(pattern)} : {support}\n")
        return d

    for _ in range(args.warmup):
        native()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        d_native = native()
    native_ms = (time.perf_counter() - t0) / args.steps * 1e3
    t0 = time.perf_counter()
    t_keys = 0.0
    for _ in range(args.steps):
        t = time.perf_counter()
        pats.keys(labels)
        t_keys += time.perf_counter() - t
    keys_ms = t_keys / args.steps * 1e3
    t0 = time.perf_counter()
    d_python = python()
    python_ms = (time.perf_counter() - t0) * 1e3
    parity = (d_native == d_python
              and (tmp / "native.txt").read_bytes() == (tmp / "python.txt").read_bytes())
    file_bytes = (tmp / "native.txt").stat().st_size
    log(f"bench: output of {n} patterns: native {native_ms:.2f} ms (keys {keys_ms:.2f} ms), "
        f"python {python_ms:.1f} ms, parity={parity}")
    return {
        "metric": "patterns rendered to getPatterns() + save() (patterns/s)",
        "value": round(n / (native_ms / 1e3), 1), "unit": "patterns/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(native_ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (numpy PCG64 seed 202)",
        "config": {"workload": "output_1m_1k", "patterns": n, "save_bytes": file_bytes,
                   "parallelism": "host (library renders; the search ran on one GPU)"},
        "roofline": None,
        "cpu_baseline": {"value": round(n / (python_ms / 1e3), 1), "unit": "patterns/s", "cores": 1,
                         "kind": "port",
                         "sample": "one run of the reference's Python rendering of the same result "
                                   "(cuApriori.py:537-552) in this process"},
        "keys_ms": round(keys_ms, 3), "python_ms": round(python_ms, 1),
        "parity": parity,
    }


def run_serve(args):
    """Served mining: ``--requests`` requests (default 10,000) with Poisson
    arrivals at ``--qps`` over a seeded pool of ``--pool`` databases drawn
    from the served workload's size distribution (T log-uniform in
    [10^4, 2*10^6], I uniform in [50, 2000], lam in [4, 12], Zipf s in
    [0.9, 1.2]); each request takes a database and its own minSup
    (log-uniform in [0.001, 0.05]).  Served by the continuous-batching server
    (slots with their own streams, stores resident under the HBM budget, cold
    at the start) against the same stream one request at a time -- the plain
    way (store built, mined and freed per request) and on resident stores.
    Every request's result is compared across the three (order-free digest);
    a bounded sample also against the oracle, which is the CPU baseline."""
    import numpy as np
    from cupami_amd import _hip, datagen
    from cupami_amd.serve import Request, Server, latency_summary, mine_one_at_a_time, mine_sequential_resident

    t = time.perf_counter()
    specs = datagen.serve_pool(args.pool)
    dbs = datagen.build_pool(specs)
    data = []
    for db in dbs:
        labels, cols = datagen.label_columns(db.items)
        data.append((db, cols, len(labels)))
    stream = datagen.serve_stream(args.requests, args.qps, len(dbs))
    reqs = []
    for i, (arrival, k, frac) in enumerate(stream):
        db, cols, items = data[k]
        reqs.append(Request(i, arrival, db.rows, cols, db.n_transactions, items,
                            int(np.ceil(frac * db.n_transactions)), key=k))
    pairs = sum(db.rows.size for db in dbs)
    log(f"bench: generated a pool of {len(dbs)} databases ({pairs} pairs) and {len(reqs)} requests "
        f"in {time.perf_counter() - t:.1f}s")

    ctx = _hip.context(0)
    budget = int(args.hbm_budget_gb * (1 << 30))
    # warm the server's contexts (buffers, graphs) on other requests, then
    # drop their stores: the timed run starts with no store resident
    server = Server(hbm_budget=budget, slots=args.slots)
    warm = [Request(i, 0.0, r.rows, r.cols, r.n_transactions, r.n_items, r.minsup, key=r.key)
            for i, r in enumerate(reqs[: min(64, len(reqs))])]
    server.run(warm)
    server.evict_all()
    mine_one_at_a_time(warm[:8], ctx)
    # long-lived objects (the pool, earlier workloads' data) are moved out of
    # the collector's reach, as a serving process does at start-up, so a
    # full collection does not stall the request loop
    import gc
    gc.collect()
    gc.freeze()
    t = time.perf_counter()
    served = server.run(reqs)
    served_s = time.perf_counter() - t
    stats = dict(server.stats)
    sizes = [len(b[3]) for b in server.batches]
    server.close()
    log(f"bench: served {len(reqs)} requests in {served_s:.2f}s ({stats})")
    t = time.perf_counter()
    alone = mine_one_at_a_time(reqs, ctx)
    alone_s = time.perf_counter() - t
    log(f"bench: one at a time {alone_s:.2f}s")
    t = time.perf_counter()
    resident = mine_sequential_resident(reqs, ctx)
    resident_s = time.perf_counter() - t
    gc.unfreeze()
    log(f"bench: one at a time, resident stores {resident_s:.2f}s")

    dig = [(o.patterns.digest(), a.patterns.digest(), q.patterns.digest())
           for o, a, q in zip(served, alone, resident)]
    parity = all(not o.rejected for o in served) and all(x == y == z for x, y, z in dig)
    n_pat = sum(len(o.patterns) for o in served)
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        threads = oracle.threads()
        hosts, expect, cpu_s, n_exp, sample = {}, [], 0.0, 0, 0
        for r in reqs:
            if cpu_s > args.cpu_sample_s:
                break
            if r.key not in hosts:
                hosts[r.key] = oracle.bits_build(r.rows, r.cols, r.n_transactions, r.n_items)
            t = time.perf_counter()
            e = oracle.eclat(hosts[r.key], r.minsup)
            cpu_s += time.perf_counter() - t
            n_exp += len(e)
            sample += 1
            parity = parity and served[r.rid].patterns == e
        cpu = {"value": round(n_exp / cpu_s, 1), "unit": "itemsets/s", "cores": threads, "kind": "port",
               "sample": f"the first {sample} requests, one run ({n_exp} itemsets in {cpu_s:.2f}s): "
                         "oracle/cpm_oracle.c vertical Eclat, OpenMP over prefix classes; each result "
                         "also compared with the served one"}
    log(f"bench: parity={parity}")
    lat = latency_summary(served)
    lat_alone = latency_summary(alone)
    lat_res = latency_summary(resident)
    bound_ok = lat["p99_ms"] is not None and lat["p99_ms"] <= args.p99_bound_ms
    line = {
        "metric": "serve p99 latency (ms), continuous batching", "value": lat["p99_ms"], "unit": "ms",
        "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": round(served_s * 1e3, 1),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (numpy PCG64: pool seeds 505000+i, request stream seed 505)",
        "config": {"workload": "serve", "requests": len(reqs), "qps": args.qps, "pool_databases": len(dbs),
                   "pool_pairs": pairs, "slots": args.slots, "hbm_budget_gb": args.hbm_budget_gb,
                   "patterns": n_pat, "parallelism": "single GPU"},
        "roofline": None,
        "cpu_baseline": cpu,
        "p99_bound_ms": args.p99_bound_ms, "p99_within_bound": bound_ok,
        "served": {**lat, "itemsets_per_s": round(n_pat / served_s, 1), "makespan_s": round(served_s, 3),
                   "batches": len(sizes), "max_batch": max(sizes), "mean_batch": round(sum(sizes) / len(sizes), 2),
                   "slots_busy_max": stats["busy_max"], "store_builds": stats["builds"],
                   "evictions": stats["evictions"], "resident_hits": stats["resident_hits"],
                   "rejected": sum(1 for o in served if o.rejected)},
        "one_at_a_time": {**lat_alone, "itemsets_per_s": round(n_pat / alone_s, 1), "makespan_s": round(alone_s, 3)},
        "one_at_a_time_resident": {**lat_res, "itemsets_per_s": round(n_pat / resident_s, 1),
                                   "makespan_s": round(resident_s, 3)},
        "parity": parity,
    }
    return line


if __name__ == "__main__":
    main()
