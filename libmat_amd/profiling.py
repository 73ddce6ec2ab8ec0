"""Text summaries of rocprofv3 CSV output.

    python -m libmat_amd.profiling stats DIR [TOP]
        per-kernel calls / average / total time from DIR/run_kernel_stats.csv
    python -m libmat_amd.profiling traffic FETCH_DIR WRITE_DIR KERNEL \
            [--copy-bytes N] [--json OUT]
        HBM bytes per dispatch of the kernels whose name contains KERNEL,
        from two `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs
        (DIR/run_counter_collection.csv), each scaled by what the same
        counter reads for the plain copy kernel k_calib_copy of N bytes
        (read N, write N).
"""
import argparse
import collections
import csv
import json
import os
import sys


def _rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def kernel_stats(d, top=16):
    rows = _rows(os.path.join(d, "run_kernel_stats.csv"))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = ["| total ms | calls | avg us | min us | max us | % | kernel |",
           "|---:|---:|---:|---:|---:|---:|---|"]
    for r in rows[:top]:
        out.append("| %.3f | %s | %.1f | %.1f | %.1f | %.1f | `%s` |" % (
            float(r["TotalDurationNs"]) / 1e6, r["Calls"],
            float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3,
            float(r["MaxNs"]) / 1e3, float(r["Percentage"]), r["Name"][:110]))
    return "\n".join(out)


def _counter(d):
    agg = collections.defaultdict(list)
    for r in _rows(os.path.join(d, "run_counter_collection.csv")):
        agg[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return agg


def traffic(fetch_dir, write_dir, kernel, copy_bytes=1 << 30):
    fe, wr = _counter(fetch_dir), _counter(write_dir)
    copy = [k for k in fe if "k_calib_copy" in k]
    if not copy:
        raise SystemExit("no k_calib_copy dispatch in the FETCH_SIZE run")
    copy = copy[0]
    sf = copy_bytes / (fe[copy][0] * 1024.0)
    sw = copy_bytes / (wr[copy][0] * 1024.0)
    name = [k for k in fe if kernel in k]
    if not name:
        raise SystemExit(f"no dispatch of a kernel named like {kernel!r}")
    name = name[0]
    f = sum(fe[name]) / len(fe[name]) * 1024.0 * sf
    w = sum(wr[name]) / len(wr[name]) * 1024.0 * sw
    lines = ["| MB read | MB written | dispatches | kernel |",
             "|---:|---:|---:|---|"]
    for k in sorted(fe, key=lambda k: -sum(fe[k]) / len(fe[k]))[:12]:
        kf = sum(fe[k]) / len(fe[k]) * 1024.0 * sf
        kw = (sum(wr.get(k, [0.0])) / max(1, len(wr.get(k, [0.0])))
              * 1024.0 * sw)
        lines.append("| %.1f | %.1f | %d | `%s` |" % (kf / 1e6, kw / 1e6,
                                                     len(fe[k]), k[:100]))
    res = {"kernel": name, "fetch_bytes": f, "write_bytes": w,
           "bytes_per_launch": f + w, "calib_copy_bytes": copy_bytes,
           "calib_fetch_kib": fe[copy][0], "calib_write_kib": wr[copy][0],
           "calib_scale_fetch": sf, "calib_scale_write": sw}
    return res, "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a1 = sub.add_parser("stats")
    a1.add_argument("dir")
    a1.add_argument("top", nargs="?", type=int, default=16)
    a2 = sub.add_parser("traffic")
    a2.add_argument("fetch_dir")
    a2.add_argument("write_dir")
    a2.add_argument("kernel")
    a2.add_argument("--copy-bytes", type=int, default=1 << 30)
    a2.add_argument("--json")
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        print(kernel_stats(a.dir, a.top))
        return
    res, table = traffic(a.fetch_dir, a.write_dir, a.kernel, a.copy_bytes)
    print(table)
    print(json.dumps(res))
    if a.json:
        with open(a.json, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
