"""Benchmark of the nsPipe time step on one MI355X.

Workload (headline): Re_bulk = 11700 pipe of length 10 pi R, NR = 96 radial
points, NT = 384 azimuthal and m_z = 1536 axial modes (reference NZ = 1537,
so the axial grid has 3(NZ-1) = 4608 points and the azimuthal grid 576),
dt = 0.001, synthetic seeded divergence-free noise on the flux-matched
Poiseuille profile.  A "step" is one time step of the reference integrator:
one nonlinear evaluation + predictor, then corrector sweeps (each a nonlinear
evaluation + solve) until the convergence measure drops below tol.

Prints one JSON line (rank 0).  value = time steps/s of the headline form,
each step one hipGraph launch (the library default).  The same number of
steps is then run kernel by kernel with HIP events recorded on the stream the
kernels run on; the roofline object is for the kernel with the largest share
of that run, from its average launch duration.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nspipe-gpu_amd"))

METRIC = "time steps/s at 96×384×1536: mixed vs FP64 GPU, % HBM roofline"
HBM_PEAK_GBS = 8000.0
CONFIG3 = dict(nr=96, nt=384, nz=1537, lz=10 * math.pi, re=11700.0, dt=0.001)
# largest configuration: Re_bulk 19000 pipe of length 4 pi R, NR=160, NT=768, m_z=3072
CONFIG5 = dict(nr=160, nt=768, nz=3073, lz=4 * math.pi, re=19000.0, dt=0.0005)
SEED5 = 20261202 + 5 * 1000
SEED3 = 20261202 + 3 * 1000
RMS3 = 0.15 * 0.5  # 0.15 U_bulk in code units (U_bulk = Qdes = 0.5)


def algorithmic_bytes(name, cfg, fp64):
    """Minimum HBM bytes of one launch under the stage byte model of the
    reference's method: every stage input read once, every output written
    once, FP64 state (16 B per complex value), transform intermediates and the
    stored nonlinear term at the precision the form stores them."""
    nr, nt, nz = cfg["nr"], cfg["nt"], cfg["nz"]
    N = nr * nt * nz
    M = nr * (3 * nt // 2) * nz
    cst, ctr, cnl = 16, (16 if fp64 else 8), (16 if fp64 else 8)
    table = {
        "dr3_kernel": 3 * N * cst + 3 * N * cnl,         # outside the stage model (d_r u at N(u)'s precision)
        "theta_inv_kernel": 3 * N * cst + 9 * M * ctr,
        "zpass_kernel": 12 * M * ctr,
        "theta_fwd_kernel": 3 * M * ctr + 3 * N * cnl,
        "linear_sweep_kernel(predictor)": 9 * N * cst + 3 * N * cnl,
        "linear_sweep_kernel(corrector)": 9 * N * cst + 3 * N * cnl,
    }
    return table.get(name)


def step_bytes(cfg, fp64, n_it):
    """B_step of the stage byte model for n_it correctors."""
    names = ["theta_inv_kernel", "zpass_kernel", "theta_fwd_kernel"]
    b_nl = sum(algorithmic_bytes(n, cfg, fp64) for n in names)
    b_sweep = algorithmic_bytes("linear_sweep_kernel(corrector)", cfg, fp64)
    return (1 + n_it) * b_nl + (1 + n_it) * b_sweep


def run_gpu(nsp, precision, steps, warmup, cfg, chosen=None, ic=None):
    p = nsp.Pipe(**cfg, precision=precision)
    if ic is None:
        p.seeded_ic(SEED3, 4, 4, 0, RMS3)
    else:
        p.seeded_ic(*ic)
    if chosen is not None:
        p.set_tuning(chosen)
        tuning = dict(tuned="set", chosen=chosen)
    else:
        tuning = p.autotune(reps=2)
    print("autotune %s: %s" % (precision, json.dumps(tuning)), file=sys.stderr, flush=True)
    if warmup:
        p.step(warmup)
    # timed region: whole steps as captured hipGraphs (the library default)
    p.synchronize()  # single process: the barrier is the device synchronisation
    t0 = time.perf_counter()
    n_it, err, flags = p.step(steps)
    p.synchronize()
    t1 = time.perf_counter()
    # the same number of steps launched kernel by kernel with HIP events on the
    # pipe's stream, for the per-kernel durations (events inside a captured
    # graph yield no elapsed times)
    p.profile_enable(True)
    p.synchronize()
    t2 = time.perf_counter()
    n_it2, _, flags2 = p.step(steps)
    p.synchronize()
    t3 = time.perf_counter()
    prof = p.profile_read()
    p.profile_enable(False)
    diag = p.diagnostics()
    p.close()
    secs = t1 - t0
    return dict(steps_per_s=steps / secs, ms_per_step=1e3 * secs / steps, n_it=n_it.tolist(),
                err_last=float(err[-1]), flags=int(max(flags.max(), flags2.max())), kernels=prof, diag=diag,
                timed_s=secs, tuning=tuning, events_steps_per_s=steps / (t3 - t2), events_n_it=n_it2.tolist())


CONFIG4A = dict(nr=48, nt=96, nz=513, lz=50.0, re=2000.0, dt=0.005)
CONFIG4B = dict(nr=64, nt=128, nz=1025, lz=50.0, re=2000.0, dt=0.005)
SEED4 = 20261202 + 4 * 1000


def run_ensemble(nsp, precision, steps, warmup, per_class=32):
    """Ensemble for predictability: 32 pipes at 48x96x513 and 32 at 64x128x1025
    (Re 2000, L = 50R, dt 0.005) stepped together by one nsp_ensemble.  Every
    member starts from the same seeded localized puff (Gaussian axial window of
    sigma 2.5R, i.e. width 5R, with rms 0.5 U_bulk inside the window) plus a
    member-specific seeded perturbation of rms 1e-6 U_bulk."""
    ms = [dict(CONFIG4A, precision=precision)] * per_class + [dict(CONFIG4B, precision=precision)] * per_class
    ens = nsp.Ensemble(ms)
    # domain rms = window rms * sqrt(mean of the squared window) = sqrt(sqrt(pi) sigma / L)
    puff_rms = 0.5 * 0.5 * math.sqrt(math.sqrt(math.pi) * 2.5 / CONFIG4A["lz"])
    for j, p in enumerate(ens.members):
        p.seeded_ic(SEED4, 4, 16, nsp.IC_PUFF, puff_rms)
        p.seeded_ic(SEED4 + 1 + j, 4, 4, nsp.IC_FLAT | nsp.IC_ADD, 1e-6 * 0.5)
    ens.autotune(1)
    if warmup:
        ens.step(warmup)
    ens[0].synchronize()
    t0 = time.perf_counter()
    n_it, err, flags = ens.step(steps)
    ens[0].synchronize()
    t1 = time.perf_counter()
    ens.close()
    secs = t1 - t0
    fp64 = precision == nsp.FP64
    # stage byte model summed over the members, per ensemble step
    step_b = sum(step_bytes(m, fp64, float(n_it[:, j].mean())) for j, m in enumerate(ms))
    achieved = step_b * steps / secs / 1e9
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), model_bytes_per_step=round(step_b))
    stages = None
    prof = load_profile("ensemble_stages.json")
    key = "fp64" if fp64 else "mixed"
    if prof and key in prof:
        # per stage: time per ensemble step from a rocprofv3 kernel trace of the
        # same ensemble (committed), against the members' summed stage bytes
        stages = {}
        for st, d in prof[key]["stages"].items():
            per_member = [algorithmic_bytes(st, m, fp64) for m in ms]
            launches = d["launches_per_step"]
            b = sum(per_member) * launches
            stages[st] = dict(us_per_step=round(d["us_per_step"], 1), alg_gbs=round(b / d["us_per_step"] / 1e3, 1),
                              alg_frac=round(b / d["us_per_step"] / 1e3 / HBM_PEAK_GBS, 4))
    return dict(members=len(ms), member_steps_per_s=round(len(ms) * steps / secs, 2),
                ensemble_steps_per_s=round(steps / secs, 4), steps=steps, n_it_mean=round(float(n_it.mean()), 3),
                flags=int(flags.max()), roofline=roof, stage_roofline=stages,
                config="32 x (NR=48, NT=96, NZ=513) + 32 x (NR=64, NT=128, NZ=1025), Re 2000, L 50R, dt 0.005",
                ic=f"puff seed {SEED4} (window rms 0.5 U_bulk, width 5R) + member j perturbation seed "
                   f"{SEED4}+1+j (rms 1e-6 U_bulk)")


def roofline(res, cfg, fp64, traffic):
    kern = {k: v for k, v in res["kernels"].items() if v[1] > 0}
    name = max(kern, key=lambda k: kern[k][0])
    ms, cnt = kern[name]
    avg_s = ms / cnt / 1e3
    alg = algorithmic_bytes(name, cfg, fp64)
    achieved = alg / avg_s / 1e9
    tr = None
    form = "fp64" if fp64 else "mixed"
    if traffic and form in traffic and name in traffic[form].get("kernels", {}):
        tr = round(traffic[form]["kernels"][name]["traffic"])
    return dict(bound="hbm", kernel=name, achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=tr, algorithmic_bytes=alg,
                avg_launch_ms=round(avg_s * 1e3, 4), launches=cnt,
                traffic_gbs=None if tr is None else round(tr / avg_s / 1e9, 1))


def stage_rooflines(res, cfg, fp64, traffic, valu=None):
    """Per stage: time per launch, algorithmic and measured-traffic rates and
    their fractions of the HBM peak."""
    form = "fp64" if fp64 else "mixed"
    out = {}
    for name, (ms, cnt) in res["kernels"].items():
        if not cnt:
            continue
        avg_s = ms / cnt / 1e3
        alg = algorithmic_bytes(name, cfg, fp64)
        d = dict(ms=round(avg_s * 1e3, 4), alg_gbs=round(alg / avg_s / 1e9, 1),
                 alg_frac=round(alg / avg_s / 1e9 / HBM_PEAK_GBS, 4))
        tk = (traffic or {}).get(form, {}).get("kernels", {}).get(name)
        if tk:
            d["traffic_gbs"] = round(tk["traffic"] / avg_s / 1e9, 1)
            d["traffic_frac"] = round(tk["traffic"] / avg_s / 1e9 / HBM_PEAK_GBS, 4)
        cs = [c for c in (valu or {}).get(form, {}).get("kernels", {}).values() if c["stage"] == name]
        if cs and all(c.get("valu_issue") is not None for c in cs):
            # VALU busy of the stage's kernels (rocprofv3 counters), weighted by their durations
            tot = sum(c["avg_us"] for c in cs)
            d["valu_issue_frac"] = round(sum(c["valu_issue"] * c["avg_us"] for c in cs) / tot, 4)
            d["valu_insts"] = round(sum(c["valu_insts"] for c in cs))
        out[name] = d
    return out


def load_profile(name):
    """A committed rocprofv3 reduction under profiles/ (None if absent)."""
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def cpu_baseline(cfg, n_it_gpu):
    """The oracle (host C++ FP64 restatement of the reference's kernels,
    OpenMP over radial levels and modes) on the same workload: one whole time
    step of the headline grid on every thread the process may use, measured
    end to end (after a one-off setup that is not timed).  A single-thread
    sample of the same work follows: one radial level of one nonlinear
    evaluation (the oracle's dominant part), timed on one thread, and the
    same per-level work on all threads for comparison.  The oracle's FFT is a
    plain recursive mixed-radix transform, so these rates are a floor for a
    CPU implementation, not a tuned CPU code."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import nsp
    from oracle_lib import Oracle, max_threads, set_threads
    # the headline IC, generated by the product's host code on the GPU
    p = nsp.Pipe(**cfg)
    p.seeded_ic(SEED3, 4, 4, 0, RMS3)
    state = p.get_state()
    p.close()
    threads = max_threads()
    o = Oracle(**cfg)
    o.set_state(*state)
    t0 = time.perf_counter()
    n_it, _ = o.step(1)
    t_step = time.perf_counter() - t0
    # single-thread sample: level 0 of one nonlinear evaluation
    o.set_state(*state)
    o.set_sample_levels(1)
    set_threads(1)
    t0 = time.perf_counter()
    o.nonlinear(*state)
    t_lvl1 = time.perf_counter() - t0
    # the same per-level work on every thread (one level per thread)
    set_threads(threads)
    o.set_sample_levels(min(threads, cfg["nr"]))
    t0 = time.perf_counter()
    o.nonlinear(*state)
    t_lvlT = time.perf_counter() - t0
    o.set_sample_levels(0)
    del o, state
    lv = min(threads, cfg["nr"])
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return dict(value=round(1.0 / t_step, 6), unit="time steps/s", cores=threads, kind="port",
                sample=(f"oracle (host C++ FP64, OpenMP, {threads} threads): one whole time step of the headline "
                        f"workload (NR=96, NT=384, NZ=1537; 1+{int(n_it[0])} sweeps) in {t_step:.1f} s"),
                step_seconds=round(t_step, 2), n_it=int(n_it[0]),
                nproc=os.cpu_count(), usable_cpus=usable, omp_num_threads=os.environ.get("OMP_NUM_THREADS"),
                single_core=dict(sample="one radial level of one nonlinear evaluation of the headline workload "
                                        "(12 inverse and 3 forward 576x4608 transforms plus the products), one thread",
                                 seconds=round(t_lvl1, 3),
                                 all_threads_seconds_per_level=round(t_lvlT / lv, 3),
                                 all_threads_levels=lv))


def copy_calibration(nbytes, copies=3):
    """Plain device-to-device copies of known size (hipMemcpy's blit kernel),
    the reference point for scaling the FETCH_SIZE / WRITE_SIZE counters of a
    profiled run.  Runs before the library launches anything."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed for the copy calibration")
    hip.hipMemset(src, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    for _ in range(copies):
        if hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed in the copy calibration")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(f"copy calibration: {copies} x {nbytes} bytes", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default="both")
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle CPU baseline")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run 3 device-to-device copies of BYTES (counter calibration)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the config-4 ensemble measurement")
    ap.add_argument("--workload", choices=["headline", "config5"], default="headline",
                    help="config5: only the largest configuration (mixed form), its per-stage rates")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="only the config-4 ensemble measurement (its own JSON line; for profiling)")
    ap.add_argument("--ensemble-steps", type=int, default=5)
    ap.add_argument("--tuning", metavar="FILE", help="apply saved kernel variants instead of autotuning")
    ap.add_argument("--save-tuning", metavar="FILE", help="write the chosen kernel variants")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import nsp
    nsp.lib()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    if args.workload == "config5":
        cfg = CONFIG5
        r = run_gpu(nsp, nsp.MIXED, args.steps, args.warmup, cfg, ic=(SEED5, 4, 4, nsp.IC_KOLMOGOROV, RMS3))
        n_it_mean = sum(r["n_it"]) / len(r["n_it"])
        out = {"metric": "time steps/s at 160x768x3072 (config 5), mixed form, % HBM roofline per stage",
               "value": round(r["steps_per_s"], 4), "unit": "time steps/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": round(r["ms_per_step"], 3), "dtype": "f32+f64",
               "data": "synthetic: seeded (mt19937_64, seed 20266202) divergence-free no-slip noise with a "
                       "k^-5/3 envelope on modes |m|<=4, l<=4, rms 0.15 U_bulk, on the flux-matched Poiseuille profile",
               "config": {"workload": "nsPipe time step, Re_bulk=19000 pipe L=4piR, NR=160 NT=768 m_z=3072 "
                                      "(reference NZ=3073), dt=0.0005", "form": "mixed", "n_it_mean": n_it_mean},
               "stage_roofline": stage_rooflines(r, cfg, False, None),
               "step_model_hbm_frac": round(step_bytes(cfg, False, n_it_mean) * r["steps_per_s"] / 1e9 / HBM_PEAK_GBS, 4),
               "tuning": (r["tuning"] or {}).get("chosen")}
        print(json.dumps(out), flush=True)
        return
    if args.ensemble_only:
        prec = nsp.FP64 if args.precision == "fp64" else nsp.MIXED
        ens = run_ensemble(nsp, prec, args.ensemble_steps, 1)
        ens["form"] = "fp64" if prec == nsp.FP64 else "mixed"
        print(json.dumps({"ensemble": ens}), flush=True)
        return
    cfg = CONFIG3
    traffic = load_profile("hbm_traffic.json")
    valu = load_profile("valu_counters.json")
    forms = {"fp64": nsp.FP64, "mixed": nsp.MIXED}
    want = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    saved = {}
    if args.tuning:
        with open(args.tuning) as fh:
            saved = json.load(fh)
    results = {}
    for f in want:
        try:
            results[f] = run_gpu(nsp, forms[f], args.steps, args.warmup, cfg, saved.get(f))
        except nsp.NspError as e:
            results[f] = dict(error=str(e))
    if args.save_tuning:
        with open(args.save_tuning, "w") as fh:
            json.dump({f: results[f]["tuning"]["chosen"] for f in want if "tuning" in results[f]}, fh, indent=1)
    head = "mixed" if "steps_per_s" in results.get("mixed", {}) else "fp64"
    r = results[head]
    fp64 = head == "fp64"
    n_it_mean = sum(r["n_it"]) / len(r["n_it"])
    out = {
        "metric": METRIC,
        "value": round(r["steps_per_s"], 4),
        "unit": "time steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if fp64 else "f32+f64",
        "data": ("synthetic: seeded (mt19937_64, seed 20264202) divergence-free no-slip noise on modes "
                 "|m|<=4, l<=4, rms 0.15 U_bulk, on the flux-matched Poiseuille profile"),
        "config": {"workload": "nsPipe time step, Re_bulk=11700 pipe L=10piR, NR=96 NT=384 m_z=1536 "
                               "(reference NZ=1537), dt=0.001",
                   "nr": cfg["nr"], "nt": cfg["nt"], "nz": cfg["nz"], "re": cfg["re"], "dt": cfg["dt"],
                   "form": head, "n_it_mean": n_it_mean},
        "roofline": roofline(r, cfg, fp64, traffic),
        "stage_roofline": stage_rooflines(r, cfg, fp64, traffic, valu),
        "step_model_hbm_frac": round(step_bytes(cfg, fp64, n_it_mean) * r["steps_per_s"] / 1e9 / HBM_PEAK_GBS, 4),
    }
    for f in want:
        rr = results[f]
        if "steps_per_s" in rr:
            out[f] = dict(steps_per_s=round(rr["steps_per_s"], 4), ms_per_step=round(rr["ms_per_step"], 3),
                          n_it=rr["n_it"], events_steps_per_s=round(rr["events_steps_per_s"], 4),
                          kernels_ms={k: round(v[0] / max(v[1], 1), 4)
                                                       for k, v in rr["kernels"].items() if v[1]},
                          ke=rr["diag"]["ke"], u_tau=rr["diag"]["u_tau"],
                          tuning=(rr["tuning"] or {}).get("chosen"))
        else:
            out[f] = rr
    if "steps_per_s" in results.get("mixed", {}) and "steps_per_s" in results.get("fp64", {}):
        out["mixed_over_fp64"] = round(results["mixed"]["steps_per_s"] / results["fp64"]["steps_per_s"], 4)
    if args.precision == "both" and "steps_per_s" in results.get("mixed", {}):
        # the refined mixed form (FP64 residual refinement of the FP32 solves),
        # with the mixed form's kernel variants; reported beside the headline
        try:
            rr = run_gpu(nsp, nsp.MIXED_REFINED, args.steps, args.warmup, cfg, results["mixed"]["tuning"]["chosen"])
            out["mixed_refined"] = dict(steps_per_s=round(rr["steps_per_s"], 4), ms_per_step=round(rr["ms_per_step"], 3),
                                        n_it=rr["n_it"], kernels_ms={k: round(v[0] / max(v[1], 1), 4)
                                                                     for k, v in rr["kernels"].items() if v[1]},
                                        ke=rr["diag"]["ke"], u_tau=rr["diag"]["u_tau"],
                                        over_fp64=(round(rr["steps_per_s"] / results["fp64"]["steps_per_s"], 4)
                                                   if "steps_per_s" in results.get("fp64", {}) else None))
        except nsp.NspError as e:
            out["mixed_refined"] = dict(error=str(e))
    if not args.no_ensemble:
        try:
            out["ensemble"] = run_ensemble(nsp, nsp.MIXED if head == "mixed" else nsp.FP64, args.ensemble_steps, 1)
            out["ensemble"]["form"] = head
        except nsp.NspError as e:
            out["ensemble"] = dict(error=str(e))
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(cfg, n_it_mean)
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = dict(error=str(e))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
