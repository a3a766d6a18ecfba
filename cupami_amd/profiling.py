"""Summaries of rocprofv3 runs, for placing kernels on this device's roofline.

Two kinds of summary are written as short markdown files of ``key: value``
lines:

* ``roofs`` -- the measured roofs from ``bench.py --workload roofs``: AND +
  popcount streaming rates with the data L2-, Infinity-Cache- and
  HBM-resident, and the FP4 MFMA loop rates;
* ``counters`` -- one kernel's per-launch counter averages from
  ``rocprofv3 --pmc`` runs of a workload: L2 requests and hits (``TCC_REQ``,
  ``TCC_HIT``, 128-byte requests), reads and writes leaving L2 for the
  fabric (``TCC_EA0_RDREQ`` in 128-byte requests, ``TCC_EA0_WRREQ`` in
  64-byte ones), and ``FETCH_SIZE`` / ``WRITE_SIZE`` scaled by what they read
  for a device-to-device copy of known size in the same run.

The Infinity Cache is memory-side: L2 counters see a request leave for the
fabric but not whether the cache or HBM answered it, so fabric bytes are an
upper bound on HBM bytes (equal when the data exceeds the cache).

Each summary records the sources it measured (a digest of their code,
comments and whitespace folded); ``read_summary`` returns nothing once they
change, so a stale measurement is never reported as current.

    python -m cupami_amd.profiling counters --kernel k_support_count \\
        --runs DIR [DIR ...] --out profiles/X.md --command "..." --sources F [F ...]
    python -m cupami_amd.profiling roofs --json roofs.json --out profiles/roofs.md --command "..."
    python -m cupami_amd.profiling stats --run DIR --title T --steps N --out profiles/X.md --command "..."
"""

from __future__ import annotations

import argparse
import csv
import hashlib
import json
import re
import sys
from collections import defaultdict
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
COPY_BYTES = 256 << 20   # the calibration copy of bench.py --copy-calibration
_TOKEN = re.compile(r'"(?:\\.|[^"\\])*"|\'(?:\\.|[^\'\\])*\'|//[^\n]*|/\*.*?\*/', re.S)


def sources_digest(files):
    """SHA-256 over the named sources (paths relative to the repository),
    comments removed and whitespace runs folded."""
    h = hashlib.sha256()
    for name in sorted(files):
        text = (ROOT / name).read_text()
        text = _TOKEN.sub(lambda m: m.group(0) if m.group(0)[0] in "\"'" else " ", text)
        h.update(name.encode())
        h.update(" ".join(text.split()).encode())
    return h.hexdigest()


def read_summary(path):
    """The ``key: value`` fields of a summary, or {} when it is missing or
    its sources have changed since it was measured."""
    path = Path(path)
    if not path.exists():
        return {}
    fields = {}
    for line in path.read_text().splitlines():
        m = re.match(r"^([a-z0-9_]+): (.*)$", line)
        if m:
            fields[m.group(1)] = m.group(2).strip()
    files = fields.get("sources", "").split()
    if files:
        try:
            if fields.get("sources_sha256") != sources_digest(files):
                return {}
        except OSError:
            return {}
    return fields


def _dispatches(run_dir):
    """{dispatch id: (kernel name, {counter: value})} of one --pmc run."""
    out = {}
    for path in Path(run_dir).rglob("*counter_collection.csv"):
        with path.open() as f:
            for r in csv.DictReader(f):
                d = int(r["Dispatch_Id"])
                name, counters = out.setdefault(d, (r["Kernel_Name"], {}))
                counters[r["Counter_Name"]] = float(r["Counter_Value"])
    return out


# one dispatch per search attempt: the level-1 setup kernel
_SEARCH = re.compile(r"\bk_level1\(")


def counter_averages(run_dirs, kernel):
    """Per-launch averages of every counter over the dispatches of
    ``kernel`` (a name fragment), the calibration copy's readings, and each
    counter's total per search in the runs that also traced ``k_level1``.

    ``k_level1`` starts every attempt of a search.  A search whose buffers
    overflow stops and starts again with larger ones (a fresh context's first
    search always does once, after level 1), so an attempt is counted as a
    search only when it made as many launches of the kernel as the run's
    longest attempt (every search of a run mines the same input); the
    attempts that stopped early are left out of the totals.  A search's
    launches past its last level exit at once: they dilute the per-launch
    averages but not the per-search totals."""
    sums, counts, copy, per_search = defaultdict(float), defaultdict(int), {}, {}
    for run in run_dirs:
        attempts = []   # per attempt: {counter: total} and the kernel's launch count
        for _, (name, counters) in sorted(_dispatches(run).items()):
            if kernel in name:
                for c, v in counters.items():
                    sums[c] += v
                    counts[c] += 1
                if attempts:
                    tot, n = attempts[-1]
                    for c, v in counters.items():
                        tot[c] += v
                    attempts[-1] = (tot, n + 1)
            elif "copyBuffer" in name:
                for c, v in counters.items():
                    copy[c] = max(copy.get(c, 0.0), v)   # the largest copy is the calibration
            elif _SEARCH.search(name):
                attempts.append((defaultdict(float), 0))
        if attempts:
            longest = max(n for _, n in attempts)
            full = [tot for tot, n in attempts if n == longest and n > 0]
            for c in full[0] if full else ():
                per_search[c] = (sum(t[c] for t in full) / len(full), len(full))
    return {c: sums[c] / counts[c] for c in sums}, counts, copy, per_search


def counters_summary(kernel, run_dirs, command, sources, workload):
    avg, counts, copy, per_search = counter_averages(run_dirs, kernel)
    lines = [f"# {kernel} counters, {workload}", "",
             "Per-launch averages over every dispatch of the kernel in the runs below (rocprofv3 --pmc),",
             "and totals per search (over the searches of each run that ran to their last level;",
             "launches past a search's last level exit at once and only dilute the per-launch averages).",
             "", "```", command, "```", ""]
    f = {"kernel": kernel, "workload": workload, "sources": " ".join(sources),
         "sources_sha256": sources_digest(sources)}
    for c, v in sorted(avg.items()):
        f[f"avg_{c.lower()}"] = f"{v:.1f}"
        f[f"dispatches_{c.lower()}"] = str(counts[c])
    if "TCC_REQ_sum" in avg:
        f["l2_request_bytes_per_launch"] = f"{avg['TCC_REQ_sum'] * 128:.0f}"
    if "TCC_HIT_sum" in avg and avg.get("TCC_REQ_sum"):
        f["l2_hit_rate"] = f"{avg['TCC_HIT_sum'] / avg['TCC_REQ_sum']:.4f}"
    if "TCC_EA0_RDREQ_sum" in avg:
        f["fabric_read_bytes_per_launch"] = f"{avg['TCC_EA0_RDREQ_sum'] * 128:.0f}"
        if copy.get("TCC_EA0_RDREQ_sum"):
            f["copy_check_rdreq_bytes"] = f"{copy['TCC_EA0_RDREQ_sum'] * 128:.0f} of {COPY_BYTES}"
    if "TCC_EA0_WRREQ_sum" in avg:
        f["fabric_write_bytes_per_launch"] = f"{avg['TCC_EA0_WRREQ_sum'] * 64:.0f}"
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        if c in avg and copy.get(c):
            scale = COPY_BYTES / (copy[c] * 1024)
            f[f"{c.lower()}_scale"] = f"{scale:.3f}"
            f[f"{c.lower()}_bytes_per_launch"] = f"{avg[c] * 1024 * scale:.0f}"
    if "fetch_size_bytes_per_launch" in f:
        f["hbm_bytes_per_launch"] = str(int(float(f["fetch_size_bytes_per_launch"])) +
                                        int(float(f.get("write_size_bytes_per_launch", 0))))
    if per_search:
        f["searches_per_run"] = " ".join(f"{c.lower()}={n}" for c, (_, n) in sorted(per_search.items()))
        tot = {c: v for c, (v, _) in per_search.items()}
        if "TCC_REQ_sum" in tot:
            f["l2_request_bytes_per_search"] = f"{tot['TCC_REQ_sum'] * 128:.0f}"
        if "TCC_EA0_RDREQ_sum" in tot:
            f["fabric_read_bytes_per_search"] = f"{tot['TCC_EA0_RDREQ_sum'] * 128:.0f}"
        if "TCC_EA0_WRREQ_sum" in tot:
            f["fabric_write_bytes_per_search"] = f"{tot['TCC_EA0_WRREQ_sum'] * 64:.0f}"
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            if c in tot and copy.get(c):
                scale = COPY_BYTES / (copy[c] * 1024)
                f[f"{c.lower()}_bytes_per_search"] = f"{tot[c] * 1024 * scale:.0f}"
        if "fetch_size_bytes_per_search" in f:
            f["hbm_bytes_per_search"] = str(int(float(f["fetch_size_bytes_per_search"])) +
                                            int(float(f.get("write_size_bytes_per_search", 0))))
    lines += [f"{k}: {v}" for k, v in f.items()]
    return "\n".join(lines) + "\n"


def roofs_summary(result, command):
    roofs = result["roofs"]
    lines = ["# Measured roofs of this device", "",
             "AND + popcount over the two halves of a buffer, best of three timed launches; XCD-sliced buffers "
             "give each XCD its own eighth (the support kernel's tile == XCD mapping), so 2-4 MiB per XCD stay "
             "in L2 and 16 MiB per XCD in the Infinity Cache.  FP4 MFMA loops of independent accumulators.  "
             "Phase boundaries: 64 rounds of 4 KB of stores per workgroup and a grid-wide barrier in one "
             "cooperative launch, against the same rounds as 64 launches of one hipGraph (best of three).",
             "", "```", command, "```", ""]
    sources = ["cupami_amd/csrc/roofs.hip", "cupami_amd/csrc/phases.hip"]
    f = {"device": result.get("device", ""), "cu_count": result.get("cu_count", ""),
         "sources": " ".join(sources), "sources_sha256": sources_digest(sources)}
    for name, r in roofs.items():
        if "us_per_barrier" in r:
            f[f"{name}_us_per_barrier"] = r["us_per_barrier"]
            f[f"{name}_us_per_launch"] = r["us_per_launch"]
        else:
            f[f"{name}_{'tops' if 'tops' in r else 'gb_s'}"] = r.get("tops", r.get("gb_s"))
    f["l2_gb_s"] = min(roofs["l2_2mib_per_xcd"]["gb_s"], roofs["l2_4mib_per_xcd"]["gb_s"])
    f["mall_gb_s"] = roofs["mall_16mib_per_xcd"]["gb_s"]
    f["hbm_gb_s"] = max(roofs["hbm_4gib"]["gb_s"], roofs["hbm_8gib_flat"]["gb_s"])
    lines += [f"{k}: {v}" for k, v in f.items()]
    return "\n".join(lines) + "\n"


def _short(name):
    name = name.replace("cpm::(anonymous namespace)::", "").replace("cpm::", "").replace("void ", "")
    return name.split("(")[0]


def stats_summary(run_dir, title, command, steps, note=""):
    """Per-kernel time of a ``--kernel-trace`` run over ``steps`` steps,
    and the launch times of the last search's support kernel, in order."""
    rows = []
    for path in Path(run_dir).rglob("*kernel_trace.csv"):
        with path.open() as f:
            rows += list(csv.DictReader(f))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    agg = defaultdict(lambda: [0, 0.0])
    for r in rows:
        a = agg[_short(r["Kernel_Name"])]
        a[0] += 1
        a[1] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
    total = sum(v[1] for v in agg.values()) or 1.0
    lines = [f"# {title}", "", note, "", "```", command, "```", "",
             "| kernel | calls | avg us | total ms | ms / step | % |", "|---|---|---|---|---|---|"]
    for k, (n, us) in sorted(agg.items(), key=lambda x: -x[1][1]):
        lines.append(f"| {k} | {n} | {us / n:.1f} | {us / 1e3:.3f} | {us / 1e3 / steps:.3f} | {100 * us / total:.2f} |")
    starts = [i for i, r in enumerate(rows) if _short(r["Kernel_Name"]) == "k_level1"]
    if starts:
        last = rows[starts[-1]:]
        sup = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in last
               if _short(r["Kernel_Name"]).startswith("k_support")]
        span = (int(last[-1]["End_Timestamp"]) - int(last[0]["Start_Timestamp"])) / 1e3
        busy = [x for x in sup if x > 5.0]
        lines += ["", f"Last search: {span:.0f} us from its first to its last kernel; support launches in order "
                      f"(us): {', '.join(f'{x:.0f}' for x in sup)}."]
        if busy:
            lines.append(f"Support launches that found work (over 5 us): {len(busy)}, mean {sum(busy) / len(busy):.1f} "
                         "us (the launches bench.py's in-kernel clock averages; launches after the last level "
                         "exit at once and are not counted there).")
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="what", required=True)
    c = sub.add_parser("counters")
    c.add_argument("--kernel", required=True)
    c.add_argument("--workload", required=True)
    c.add_argument("--runs", nargs="+", required=True)
    c.add_argument("--sources", nargs="+", required=True)
    c.add_argument("--command", required=True)
    c.add_argument("--out", required=True)
    t = sub.add_parser("stats")
    t.add_argument("--run", required=True)
    t.add_argument("--title", required=True)
    t.add_argument("--steps", type=int, required=True)
    t.add_argument("--note", default="")
    t.add_argument("--command", required=True)
    t.add_argument("--out", required=True)
    r = sub.add_parser("roofs")
    r.add_argument("--json", required=True)
    r.add_argument("--command", required=True)
    r.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    if a.what == "counters":
        text = counters_summary(a.kernel, a.runs, a.command, a.sources, a.workload)
    elif a.what == "stats":
        text = stats_summary(a.run, a.title, a.command, a.steps, a.note)
    else:
        text = roofs_summary(json.loads(Path(a.json).read_text().strip().splitlines()[-1]), a.command)
    Path(a.out).write_text(text)
    print(text)


if __name__ == "__main__":
    sys.exit(main())
