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


def run_gpu(nsp, precision, steps, warmup, cfg, chosen=None):
    p = nsp.Pipe(**cfg, precision=precision)
    p.seeded_ic(SEED3, 4, 4, 0, RMS3)
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
    t0 = time.perf_counter()
    n_it, err, flags = ens.step(steps)
    t1 = time.perf_counter()
    ens.close()
    secs = t1 - t0
    return dict(members=len(ms), member_steps_per_s=round(len(ms) * steps / secs, 2),
                ensemble_steps_per_s=round(steps / secs, 4), steps=steps, n_it_mean=round(float(n_it.mean()), 3),
                flags=int(flags.max()),
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


def stage_rooflines(res, cfg, fp64, traffic):
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
        out[name] = d
    return out


def load_traffic():
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def cpu_baseline(cfg, n_it_gpu):
    """Oracle on a bounded sample: the headline workload with NR reduced to
    8 (same 576 x 4608 transforms per level, same per-mode solves), one
    time step, scaled linearly in NR and by sweep count to one full step."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import nsp
    from oracle_lib import Oracle, max_threads
    nr_s = 8
    sample = dict(cfg, nr=nr_s)
    # the same IC recipe on the sample grid, generated by the product's host code
    p = nsp.Pipe(**sample)
    p.seeded_ic(SEED3, 4, 4, 0, RMS3)
    state = p.get_state()
    p.close()
    o = Oracle(**sample)
    o.set_state(*state)
    t0 = time.perf_counter()
    n_it, _ = o.step(1)
    t = time.perf_counter() - t0
    per_sweep = t / (1 + int(n_it[0])) * (cfg["nr"] / nr_s)
    full_step = per_sweep * (1 + n_it_gpu)
    return dict(value=round(1.0 / full_step, 6), unit="time steps/s", cores=max_threads(), kind="port",
                sample=(f"oracle (host C++ FP64, OpenMP) one time step of the headline workload with NR={nr_s} "
                        f"(same 576x4608 transforms and per-mode solves), {t:.1f} s for 1+{int(n_it[0])} sweeps, "
                        f"scaled x{cfg['nr'] // nr_s} in NR and to 1+{n_it_gpu:.2f} sweeps"),
                sample_seconds=round(t, 2))


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
    cfg = CONFIG3
    traffic = load_traffic()
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
        "stage_roofline": stage_rooflines(r, cfg, fp64, traffic),
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
