"""Short text summaries of rocprofv3 CSV output (for the profiles/ folder).

    python -m orac_hip.rocprof_summary stats <kernel_stats.csv> <out.txt> "<profiled command>"
    python -m orac_hip.rocprof_summary pmc <counter_collection.csv> <out.txt> "<profiled command>"
    python -m orac_hip.rocprof_summary traffic <fetch.csv> <write.csv> "<kernel name part>" <probe bytes> \
        <out.txt> "<FETCH_SIZE command>" "<WRITE_SIZE command>"

``stats`` lists every kernel's calls and average/min/max/total duration from
``rocprofv3 --kernel-trace --stats --output-format csv``.  ``pmc`` averages
each counter per dispatch and kernel from ``rocprofv3 --pmc ... --output-format
csv`` (FETCH_SIZE and WRITE_SIZE are in KiB per dispatch).  ``traffic`` turns
the FETCH_SIZE and WRITE_SIZE runs of one kernel into HBM bytes per launch,
scaled by the counters' reading on k_probe_copy.  Column names are
matched loosely, so minor CSV layout changes between rocprofv3 versions do
not matter.
"""
from __future__ import annotations

import csv
import hashlib
import os
import sys
from collections import defaultdict


def _col(header, *keys, exclude=()):
    low = [h.lower() for h in header]
    for k in keys:
        for i, h in enumerate(low):
            if k in h and not any(x in h for x in exclude):
                return i
    raise KeyError(f"no column like {keys} in {header}")


def stats(path, out, cmd):
    with open(path, newline="") as f:
        rows = list(csv.reader(f))
    h, body = rows[0], rows[1:]
    iname = _col(h, "name")
    icalls = _col(h, "calls")
    itot = _col(h, "totaldurationns", "total")
    iavg = _col(h, "averagens", "average")
    imin = _col(h, "minns", "min")
    imax = _col(h, "maxns", "max")
    body.sort(key=lambda r: -float(r[itot]))
    tot = sum(float(r[itot]) for r in body) or 1.0
    with open(out, "w") as f:
        f.write(f"command: {cmd}\n")
        f.write("per-kernel durations from rocprofv3 --kernel-trace --stats, microseconds\n\n")
        f.write(f"{'calls':>6} {'avg_us':>9} {'min_us':>9} {'max_us':>9} {'total_us':>10} {'pct':>6}  kernel\n")
        for r in body:
            t = float(r[itot])
            f.write(f"{int(float(r[icalls])):6d} {float(r[iavg]) / 1e3:9.2f} {float(r[imin]) / 1e3:9.2f} "
                    f"{float(r[imax]) / 1e3:9.2f} {t / 1e3:10.1f} {100 * t / tot:6.2f}  {r[iname][:110]}\n")


def pmc(path, out, cmd):
    with open(path, newline="") as f:
        rows = list(csv.reader(f))
    h, body = rows[0], rows[1:]
    iname = _col(h, "kernel_name", "kernel-name", "name", exclude=("counter",))
    icn = _col(h, "counter_name", "counter-name")
    icv = _col(h, "counter_value", "counter-value", "value")
    idisp = None
    try:
        idisp = _col(h, "dispatch_id", "dispatch-id", "correlation")
    except KeyError:
        pass
    # sum over the per-dimension rows of one dispatch, then average over dispatches
    per = defaultdict(float)
    for r in body:
        key = (r[iname], r[icn], r[idisp] if idisp is not None else len(per))
        per[key] += float(r[icv])
    agg = defaultdict(list)
    for (name, cn, _), v in per.items():
        agg[(name, cn)].append(v)
    with open(out, "w") as f:
        f.write(f"command: {cmd}\n")
        f.write("per-kernel counter averages per dispatch from rocprofv3 --pmc (FETCH_SIZE / WRITE_SIZE in KiB)\n\n")
        f.write(f"{'dispatches':>10} {'average':>14}  counter  kernel\n")
        for (name, cn), vals in sorted(agg.items(), key=lambda kv: -sum(kv[1]) / len(kv[1])):
            f.write(f"{len(vals):10d} {sum(vals) / len(vals):14.1f}  {cn:10s} {name[:100]}\n")


def _pmc_avg(path, kernel_part, counter):
    """Average per dispatch of `counter` for the kernels whose name contains kernel_part."""
    with open(path, newline="") as f:
        rows = list(csv.reader(f))
    h, body = rows[0], rows[1:]
    iname = _col(h, "kernel_name", "kernel-name", "name", exclude=("counter",))
    icn = _col(h, "counter_name", "counter-name")
    icv = _col(h, "counter_value", "counter-value", "value")
    idisp = _col(h, "dispatch_id", "dispatch-id", "correlation")
    per = defaultdict(float)
    for r in body:
        if kernel_part in r[iname] and r[icn] == counter:
            per[r[idisp]] += float(r[icv])
    if not per:
        raise KeyError(f"no {counter} rows for a kernel like {kernel_part!r} in {path}")
    return sum(per.values()) / len(per), len(per)


def traffic(fetch_csv, write_csv, kernel, probe_bytes, out, cmd_fetch, cmd_write):
    """HBM bytes per launch of `kernel`: FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch) each scaled by what the counter reads for k_probe_copy, a plain
    copy kernel that reads and writes probe_bytes."""
    f_k, nf = _pmc_avg(fetch_csv, kernel, "FETCH_SIZE")
    w_k, nw = _pmc_avg(write_csv, kernel, "WRITE_SIZE")
    f_p, _ = _pmc_avg(fetch_csv, "k_probe_copy", "FETCH_SIZE")
    w_p, _ = _pmc_avg(write_csv, "k_probe_copy", "WRITE_SIZE")
    fs = probe_bytes / (f_p * 1024.0)
    ws = probe_bytes / (w_p * 1024.0)
    fetched, written = f_k * 1024.0 * fs, w_k * 1024.0 * ws
    with open(out, "w") as f:
        f.write(f"kernel_source_sha16: {source_sha16()}\n")
        f.write(f"command (FETCH_SIZE): {cmd_fetch}\n")
        f.write(f"command (WRITE_SIZE): {cmd_write}\n")
        f.write("HBM bytes per launch from FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled by the\n")
        f.write("same counter on k_probe_copy, a plain copy kernel reading and writing probe_copy_bytes\n\n")
        f.write(f"kernel: {kernel}\n")
        f.write(f"dispatches: {nf} (FETCH_SIZE), {nw} (WRITE_SIZE)\n")
        f.write(f"fetch_kib_per_dispatch: {f_k:.1f}\n")
        f.write(f"write_kib_per_dispatch: {w_k:.1f}\n")
        f.write(f"probe_copy_bytes: {probe_bytes}\n")
        f.write(f"probe_fetch_kib: {f_p:.1f}\n")
        f.write(f"probe_write_kib: {w_p:.1f}\n")
        f.write(f"fetch_scale: {fs:.4f}\n")
        f.write(f"write_scale: {ws:.4f}\n")
        f.write(f"fetched_bytes_per_launch: {fetched:.0f}\n")
        f.write(f"written_bytes_per_launch: {written:.0f}\n")
        f.write(f"traffic_bytes_per_launch: {fetched + written:.0f}\n")


PAIR_SOURCE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "csrc", "force_pairs.hip")


def source_sha16(path=PAIR_SOURCE):
    """First 16 hex digits of the SHA-256 of the pair kernel's source file."""
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def read_stamp(path):
    """Which source a committed traffic summary was measured on, and whether
    that is still the current pair kernel source."""
    stamp = None
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("kernel_source_sha16:"):
                    stamp = line.split(":", 1)[1].strip()
    except OSError:
        return {"profile": os.path.basename(path), "kernel_source_sha16": None, "current": False}
    return {"profile": os.path.basename(path), "kernel_source_sha16": stamp, "current": stamp == source_sha16()}


def read_traffic(path):
    """traffic_bytes_per_launch from a summary written by traffic(), or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def main(argv):
    if len(argv) == 9 and argv[1] == "traffic":
        traffic(argv[2], argv[3], argv[4], int(argv[5]), argv[6], argv[7], argv[8])
        print(open(argv[6]).read())
        return
    if len(argv) != 5 or argv[1] not in ("stats", "pmc"):
        sys.exit(__doc__)
    (stats if argv[1] == "stats" else pmc)(argv[2], argv[3], argv[4])
    print(open(argv[3]).read())


if __name__ == "__main__":
    main(sys.argv)
