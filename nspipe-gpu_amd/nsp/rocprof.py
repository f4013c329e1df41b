"""Reduce rocprofv3 output of a bench run to per-stage figures.

Kernel names are grouped into the stages the library's own event profiler
reports (nsp_profile_name): dr3_kernel, theta_inv_kernel, zpass_kernel,
theta_fwd_kernel, linear_sweep_kernel(predictor|corrector).

HBM traffic comes from two counter runs of the same command (FETCH_SIZE, then
WRITE_SIZE; rocprofv3 reports both in KiB per dispatch).  Each counter is
divided by what it reads for a plain device-to-device copy of known size
(hipMemcpy's blit kernel), run in the same process before the library does
any work: bytes = counter_KiB * 1024 / (counter_KiB_of_copy * 1024 / copy_bytes).

    python -m nsp.rocprof traffic --fetch F.csv --write W.csv --copy-bytes B --form mixed \
        --out profiles/hbm_traffic.json
    python -m nsp.rocprof stats --stats S_kernel_stats.csv --title "..." --command "..." \
        --out profiles/x.txt
"""
import argparse
import csv
import json
import os
import re
import sys

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def stage_of(name):
    """Library stage of a rocprof kernel name, or None for other kernels."""
    n = name
    if "dr3_kernel" in n:
        return "dr3_kernel"
    if "theta_inv" in n:
        return "theta_inv_kernel"
    if "zpass" in n:
        return "zpass_kernel"
    if "theta_fwd" in n:
        return "theta_fwd_kernel"
    m = re.search(r"sweep_kernel<(\d)>", n)
    if m:
        return "linear_sweep_kernel(predictor)" if m.group(1) == "1" else "linear_sweep_kernel(corrector)"
    return None


def read_counters(path, counter):
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    return rows


def calibrate(rows, copy_bytes):
    """KiB the counter reads per byte of the calibration copy: the copy kernels
    dispatched before the first library kernel."""
    total, n = 0.0, 0
    for _, name, v in rows:
        if name.startswith("nsp::") or "nsp::" in name:
            break
        if COPY_KERNEL in name:
            total += v
            n += 1
    if n == 0:
        raise SystemExit("no calibration copy dispatches found before the first library kernel")
    return total * 1024.0, n


def traffic(args):
    fetch = read_counters(args.fetch, "FETCH_SIZE")
    write = read_counters(args.write, "WRITE_SIZE")
    fb, nf = calibrate(fetch, args.copy_bytes)
    wb, nw = calibrate(write, args.copy_bytes)
    # the calibration copies ran nf (nw) times over copy_bytes each
    ffac = fb / (nf * args.copy_bytes) if args.copies is None else fb / (args.copies * args.copy_bytes)
    wfac = wb / (nw * args.copy_bytes) if args.copies is None else wb / (args.copies * args.copy_bytes)
    out = {}
    if os.path.exists(args.out):
        with open(args.out) as f:
            out = json.load(f)
    per = {}
    for rows, key, fac in ((fetch, "fetch_bytes", ffac), (write, "write_bytes", wfac)):
        acc = {}
        for _, name, v in rows:
            st = stage_of(name)
            if st is None:
                continue
            a = acc.setdefault(st, [0.0, 0])
            a[0] += v * 1024.0 / fac
            a[1] += 1
        for st, (tot, cnt) in acc.items():
            per.setdefault(st, {})[key] = tot / cnt
            per[st]["dispatches"] = cnt
    for st, d in per.items():
        d["traffic"] = d.get("fetch_bytes", 0.0) + d.get("write_bytes", 0.0)
    out[args.form] = dict(calibration=dict(copy_bytes=args.copy_bytes, fetch_counter_per_byte=ffac,
                                           write_counter_per_byte=wfac),
                          kernels=per, command=args.command or "")
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out[args.form], indent=1))


def stats(args):
    rows = []
    with open(args.stats) as f:
        for r in csv.DictReader(f):
            rows.append(r)
    lines = [f"# {args.title}", f"# command: {args.command}",
             "# source: rocprofv3 --kernel-trace --stats, kernel_stats.csv (durations in microseconds)",
             f"{'kernel':70s} {'stage':32s} {'calls':>6s} {'total_us':>12s} {'avg_us':>10s} {'pct':>6s}"]
    for r in rows:
        name = r["Name"]
        st = stage_of(name) or "-"
        short = re.sub(r"\(.*$", "", name).replace("nsp::", "")
        calls = int(r["Calls"])
        tot = float(r["TotalDurationNs"]) / 1e3
        avg = float(r["AverageNs"]) / 1e3
        pct = float(r["Percentage"])
        lines.append(f"{short[:70]:70s} {st:32s} {calls:6d} {tot:12.1f} {avg:10.1f} {pct:6.2f}")
    text = "\n".join(lines) + "\n"
    with open(args.out, "w") as f:
        f.write(text)
    sys.stdout.write(text)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m nsp.rocprof")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--copies", type=int, default=None, help="calibration copies (default: dispatch count)")
    t.add_argument("--form", required=True)
    t.add_argument("--command", default="")
    t.add_argument("--out", required=True)
    s = sub.add_parser("stats")
    s.add_argument("--stats", required=True)
    s.add_argument("--title", required=True)
    s.add_argument("--command", required=True)
    s.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    (traffic if args.cmd == "traffic" else stats)(args)


if __name__ == "__main__":
    main()
