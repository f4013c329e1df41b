"""Reduce rocprofv3 output of a bench run to per-stage figures.

Kernel names are grouped into the stages the library's own event profiler
reports (nsp_profile_name): dr3_kernel, theta_inv_kernel, zpass_kernel,
theta_fwd_kernel, linear_sweep_kernel(predictor|corrector); byte figures are
per stage launch (a sweep launch is its four pass kernels).

HBM traffic comes from two counter runs of the same command (FETCH_SIZE, then
WRITE_SIZE; rocprofv3 reports both in KiB per dispatch).  Each counter is
divided by what it reads for a plain device-to-device copy of known size
(hipMemcpy's blit kernel), run in the same process before the library does
any work: bytes = counter_KiB * 1024 / (counter_KiB_of_copy * 1024 / copy_bytes).

    python -m nsp.rocprof traffic --fetch F.csv --write W.csv --copy-bytes B --form mixed \
        --out profiles/hbm_traffic.json
    python -m nsp.rocprof stats --stats S_kernel_stats.csv --trace S_kernel_trace.csv \
        --title "..." --command "..." --out profiles/x.txt
    python -m nsp.rocprof counters --pmc P_counter_collection.csv --form mixed \
        --title "..." --command "..." --out profiles/y.txt --json profiles/valu_counters.json
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
    m = re.search(r"sweep_kernel<(\d), (\d)", n)  # <pass, predictor, ...>
    if m:
        return "linear_sweep_kernel(predictor)" if m.group(2) == "1" else "linear_sweep_kernel(corrector)"
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
    of the calibration size dispatched before the first library kernel (the
    small copies of table uploads are left out: under a quarter of the largest)."""
    vals = []
    for _, name, v in rows:
        if "nsp::" in name:
            break
        if COPY_KERNEL in name:
            vals.append(v)
    big = [v for v in vals if vals and v >= 0.25 * max(vals)]
    if not big:
        raise SystemExit("no calibration copy dispatches found before the first library kernel")
    return sum(big) * 1024.0, len(big)


def stage_bytes(rows, fac):
    """Bytes per executed launch of each stage.  Launches that return at entry
    (correctors of a captured step after convergence) are left out: per kernel
    name, dispatches under 5% of its largest.  A stage launch is one dispatch of
    each of its kernels, so the per-name averages add up."""
    by_name = {}
    for _, name, v in rows:
        if stage_of(name) is not None:
            by_name.setdefault(name, []).append(v * 1024.0 / fac)
    per, count = {}, {}
    for name, vs in by_name.items():
        top = max(vs)
        kept = [v for v in vs if v >= 0.05 * top] if top > 0 else vs
        st = stage_of(name)
        per[st] = per.get(st, 0.0) + sum(kept) / len(kept)
        count[st] = count.get(st, 0) + len(kept)
    return per, count


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
        b, n = stage_bytes(rows, fac)
        for st in b:
            per.setdefault(st, {})[key] = b[st]
            per[st]["dispatches"] = n[st]
    for st, d in per.items():
        d["traffic"] = d.get("fetch_bytes", 0.0) + d.get("write_bytes", 0.0)
    out[args.form] = dict(calibration=dict(copy_bytes=args.copy_bytes, fetch_counter_per_byte=ffac,
                                           write_counter_per_byte=wfac),
                          kernels=per, command=args.command or "")
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out[args.form], indent=1))


def stats(args):
    """kernel_stats.csv as text, plus, from the kernel trace, the launches that
    did the work: correctors captured in a step graph after convergence, and
    their nonlinear evaluations, return at entry and are left out (launches
    under 5% of the kernel's longest), with the per-stage sums per launch."""
    rows = []
    with open(args.stats) as f:
        for r in csv.DictReader(f):
            rows.append(r)
    durs = {}
    if args.trace:
        with open(args.trace) as f:
            for r in csv.DictReader(f):
                d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
                durs.setdefault(r["Kernel_Name"], []).append(d)
    lines = [f"# {args.title}", f"# command: {args.command}",
             "# source: rocprofv3 --kernel-trace --stats (kernel_stats.csv, kernel_trace.csv); microseconds",
             "# exec / exec_avg_us: launches that did the work (>= 5% of the kernel's longest) and their mean",
             f"{'kernel':70s} {'stage':32s} {'calls':>6s} {'total_us':>12s} {'avg_us':>10s} {'pct':>6s} "
             f"{'exec':>5s} {'exec_avg_us':>11s}"]
    stage_sum = {}
    for r in rows:
        name = r["Name"]
        st = stage_of(name) or "-"
        short = re.sub(r"\(.*$", "", name).replace("nsp::", "")
        calls = int(r["Calls"])
        tot = float(r["TotalDurationNs"]) / 1e3
        avg = float(r["AverageNs"]) / 1e3
        pct = float(r["Percentage"])
        ds = durs.get(name, [])
        kept = [d for d in ds if ds and d >= 0.05 * max(ds)]
        ex = len(kept)
        ea = sum(kept) / ex if ex else 0.0
        if st != "-" and ex:
            stage_sum.setdefault(st, []).append((short, ea))
        lines.append(f"{short[:70]:70s} {st:32s} {calls:6d} {tot:12.1f} {avg:10.1f} {pct:6.2f} {ex:5d} {ea:11.1f}")
    if stage_sum:
        lines.append("# per stage launch (sum of its kernels' exec_avg_us; several forms of one stage")
        lines.append("# appear when the run also timed other variants):")
        for st, ks in sorted(stage_sum.items()):
            lines.append(f"#   {st:32s} {sum(a for _, a in ks):10.1f} us  ({', '.join(k for k, _ in ks)})")
    text = "\n".join(lines) + "\n"
    with open(args.out, "w") as f:
        f.write(text)
    sys.stdout.write(text)


CU_NUM = 256  # MI355X: 8 XCDs x 32 CUs
XCD_NUM = 8   # GRBM_GUI_ACTIVE is reported summed over the XCDs


def counters(args):
    """Per-stage instruction and LDS counters of one rocprofv3 --pmc run over
    the executed launches (>= 5% of the kernel's longest), averaged per
    dispatch.  With C = GRBM_GUI_ACTIVE / XCD_NUM (GPU cycles of the dispatch;
    the counter comes summed over the eight XCDs):
      valu_issue = SQ_ACTIVE_INST_VALU / (CU_NUM * C): VALU busy, the fraction
                   of the one-wave-instruction-per-CU-cycle issue peak;
      waves_per_cu = 4 * SQ_WAVE_CYCLES / (CU_NUM * C): mean resident waves
                   (SQ_WAVE_CYCLES counts in units of 4 cycles);
      lds_active = SQ_LDS_IDX_ACTIVE / (CU_NUM * C);
      lds_conflict = SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE (share of the
                   active LDS cycles lost to bank conflicts)."""
    per = {}
    meta = {}
    with open(args.pmc) as f:
        for r in csv.DictReader(f):
            d = int(r["Dispatch_Id"])
            per.setdefault(d, {})[r["Counter_Name"]] = float(r["Counter_Value"])
            meta[d] = (r["Kernel_Name"], (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3,
                       int(r["VGPR_Count"]), int(r["LDS_Block_Size"]), int(r["Workgroup_Size"]))
    by_name = {}
    for d, c in per.items():
        name, us, vgpr, lds, wg = meta[d]
        if stage_of(name) is None:
            continue
        by_name.setdefault(name, []).append((us, c, vgpr, lds, wg))
    out = {}
    lines = [f"# {args.title}", f"# command: {args.command}",
             "# source: rocprofv3 --pmc counter_collection.csv; per executed launch (>= 5% of the kernel's longest)",
             f"# cycles = GRBM_GUI_ACTIVE/{XCD_NUM} (summed over the XCDs); valu_issue = SQ_ACTIVE_INST_VALU/"
             f"({CU_NUM}*cycles); waves/CU = 4*SQ_WAVE_CYCLES/({CU_NUM}*cycles); lds_active = "
             f"SQ_LDS_IDX_ACTIVE/({CU_NUM}*cycles); lds_conflict = SQ_LDS_BANK_CONFLICT/SQ_LDS_IDX_ACTIVE",
             f"{'kernel':62s} {'stage':32s} {'n':>3s} {'us':>8s} {'VALU_M':>8s} {'valu_iss':>8s} "
             f"{'waves/CU':>8s} {'LDS_M':>7s} {'lds_act':>7s} {'lds_cfl':>7s} {'lds_B':>6s}"]
    for name, rs in sorted(by_name.items(), key=lambda kv: -max(x[0] for x in kv[1])):
        top = max(x[0] for x in rs)
        kept = [x for x in rs if x[0] >= 0.05 * top]
        n = len(kept)
        avg = lambda key: sum(x[1].get(key, 0.0) for x in kept) / n
        us = sum(x[0] for x in kept) / n
        cyc = avg("GRBM_GUI_ACTIVE") / XCD_NUM
        valu, act, wcyc = avg("SQ_INSTS_VALU"), avg("SQ_ACTIVE_INST_VALU"), avg("SQ_WAVE_CYCLES")
        lds_i, lds_a, lds_c = avg("SQ_INSTS_LDS"), avg("SQ_LDS_IDX_ACTIVE"), avg("SQ_LDS_BANK_CONFLICT")
        d = dict(stage=stage_of(name), dispatches=n, avg_us=us, gpu_cycles=cyc, valu_insts=valu,
                 waves=avg("SQ_WAVES"), valu_issue=act / (CU_NUM * cyc) if cyc else None,
                 waves_per_cu=4 * wcyc / (CU_NUM * cyc) if cyc else None, lds_insts=lds_i,
                 lds_active=lds_a / (CU_NUM * cyc) if cyc else None,
                 lds_conflict=lds_c / lds_a if lds_a else 0.0, lds_bytes=kept[0][3], workgroup=kept[0][4])
        short = re.sub(r"\(.*$", "", name).replace("nsp::", "")
        out[short] = d
        f = lambda v, w=8, p=3: f"{v:{w}.{p}f}" if v is not None else f"{'-':>{w}s}"
        lines.append(f"{short[:62]:62s} {d['stage']:32s} {n:3d} {us:8.1f} {valu / 1e6:8.1f} {f(d['valu_issue'])} "
                     f"{f(d['waves_per_cu'], 8, 2)} {lds_i / 1e6:7.1f} {f(d['lds_active'], 7)} "
                     f"{f(d['lds_conflict'], 7)} {d['lds_bytes']:6d}")
    text = "\n".join(lines) + "\n"
    with open(args.out, "w") as fh:
        fh.write(text)
    if args.json:
        js = {}
        if os.path.exists(args.json):
            with open(args.json) as fh:
                js = json.load(fh)
        js[args.form] = dict(command=args.command, kernels=out)
        with open(args.json, "w") as fh:
            json.dump(js, fh, indent=1, sort_keys=True)
    sys.stdout.write(text)


def ensemble(args):
    """Per-stage time per ensemble step from the kernel trace of a bench
    ensemble run whose last args.steps ensemble steps are the timed ones.
    Each step starts with the batched radial derivative of u^n and has one
    batched predictor pass A (sweep_kernel<0, 1, 1, ...>); the last `steps`
    of those anchor the window.  Per stage: summed duration per step and the
    number of evaluations per step (nonlinear stages: radial-derivative
    launches, one per evaluation for all members; sweeps: pass-A launches)."""
    rows = []
    with open(args.trace) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()
    pred_a = [i for i, (_, _, n) in enumerate(rows) if re.search(r"sweep_kernel<0, 1, 1", n)]
    if len(pred_a) < args.steps:
        raise SystemExit("fewer batched predictor launches than steps")
    first = pred_a[-args.steps]
    start = max(i for i in range(first) if "dr3_kernel" in rows[i][2])
    tot, cnt = {}, {}
    for t0, t1, n in rows[start:]:
        st = stage_of(n)
        if st is None:
            continue
        tot[st] = tot.get(st, 0.0) + (t1 - t0) / 1e3
        if "dr3_kernel" in n or re.search(r"sweep_kernel<0, ", n):
            cnt[st] = cnt.get(st, 0) + 1
    evals = cnt.get("dr3_kernel", 0) / args.steps
    out = {}
    if os.path.exists(args.out):
        with open(args.out) as f:
            out = json.load(f)
    stages = {}
    for st, us in tot.items():
        ev = evals if not st.startswith("linear_sweep") else cnt.get(st, 0) / args.steps
        stages[st] = dict(us_per_step=us / args.steps, launches_per_step=ev)
    out[args.form] = dict(command=args.command, steps=args.steps, stages=stages)
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out[args.form], indent=1))


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
    s.add_argument("--trace", default=None, help="kernel_trace.csv of the same run")
    s.add_argument("--title", required=True)
    s.add_argument("--command", required=True)
    s.add_argument("--out", required=True)
    c = sub.add_parser("counters")
    c.add_argument("--pmc", required=True, help="counter_collection.csv of a --pmc run")
    c.add_argument("--title", required=True)
    c.add_argument("--command", required=True)
    c.add_argument("--form", required=True)
    c.add_argument("--out", required=True)
    c.add_argument("--json", default=None, help="also merge per-kernel figures into this JSON under --form")
    e = sub.add_parser("ensemble")
    e.add_argument("--trace", required=True, help="kernel_trace.csv of a bench ensemble run")
    e.add_argument("--steps", type=int, required=True, help="timed ensemble steps at the end of the trace")
    e.add_argument("--form", required=True)
    e.add_argument("--command", required=True)
    e.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    {"traffic": traffic, "stats": stats, "counters": counters, "ensemble": ensemble}[args.cmd](args)


if __name__ == "__main__":
    main()
