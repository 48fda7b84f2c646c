#!/usr/bin/env python3
"""Benchmark of the period scan hot path: trial-period Levenberg-Marquardt fits per second at
Lmax = Mmax = 6 on 288 facets (the production shape: 40 dense lightcurves x 50 points + 400 sparse
calibrated points, ndata = 2400, 55 fitted parameters, stop condition 50 = fixed 50 LM iterations),
FP64 and mixed precision on the same fits, with the best-period agreement of the two.

A step is one ps_fit_batch call over a window of trial periods x 10 initial poles, inputs resident
in HBM (uploaded once by ps_prepare); ps_fit_batch returns after the library's stream has
synchronised, so the wall clock around the K steps covers all device work.  Step k scans the k-th
window of the grid, the first one centred on the true period.  The default run times the FP64
kernels (the line's value) and then the mixed-precision kernels on the same windows; it reports the
mixed rate, the speedup, whether both precisions pick the same best period in every window and the
per-period chi2 difference.  The roofline figures use the fit kernel's own duration, measured with
HIP events recorded on the stream it is launched on.

    python bench.py [--steps K] [--warmup W] [--precision both|fp64|mixed] [--periods P]
                    [--workload config1|config3|config4] [--batch-periods B]

The default workload (config1) is the headline; config3 (a sparse survey fitted on 512 facets, Dsph
read through L2) is a second shape, reported under its own metric name; config4 is the batched uneven
workload (512 asteroids, one graph-captured persistent launch over a device work queue, FP64).
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "trial-period LM fits/sec at Lmax=6, 288 facets: mixed vs FP64, same best period"
METRIC_CONFIG3 = "trial-period LM fits/sec at Lmax=6, 512 facets, sparse survey (5000 points)"
METRIC_CONFIG2 = "model evaluations/sec (forward dyda + J^T W J, J^T W r without storing J) at Lmax=6, 288 facets"
METRIC_CONFIG4 = "LM fits/sec of a batch of 512 asteroids (ndata 100-20000, Lmax 3/4/6, 288/512 facets), one persistent launch"
PEAK_FP64 = 78.6   # MI355X dense FP64 matrix peak, TFLOP/s
PEAK_FP32 = 157.3  # MI355X dense FP32 matrix peak, TFLOP/s
ORACLE_CLI = os.path.join(ROOT, "oracle", "ps_oracle_cli")
REF_CLI = os.path.join(ROOT, "oracle", "_ref", "ps_ref_cli")
REF_FMA_CLI = os.path.join(ROOT, "oracle", "_ref", "ps_ref_fma_cli")
# the host CPU share of one GPU on the benchmark machines: worker pools stay within it
MAX_CPU_WORKERS = 16


def flops_per_eval(info):
    """Algorithmic flops of one model evaluation, N = ndata + 3 convexity rows: the shape-Jacobian
    contraction N * 2 Nf Nc, and the normal equations N * (mfit (mfit + 1) + 2 mfit) (the J^T W J lower
    triangle with diagonal, and J^T W r)."""
    n, nf, nc, mf = info["npts"], info["nfac"], info["ncoef"], info["mfit"]
    return n * 2 * nf * nc, n * (mf * (mf + 1) + 2 * mf)


def traffic_per_launch(workload, precision, kernel):
    """HBM bytes per launch from the two counter passes committed under profiles/ (null when the kernel
    form was not measured on this workload)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}_{precision}_{kernel}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def cpu_cores():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(MAX_CPU_WORKERS, n))


def cpu_baseline(cli, label, kind, input_path, conw_r, seconds_target):
    """A single-threaded CPU build of the fit, one process per host core (at most the GPU's CPU share),
    on disjoint trial periods of the same workload -- the reference's own data-parallel split.  Each
    process runs whole periods (10 poles) for about seconds_target."""
    if not os.path.exists(cli):
        return None
    cores = cpu_cores()
    probe = subprocess.run([cli, "fits", input_path, "1", "1", repr(conw_r), "1"], capture_output=True, text=True,
                           check=True)
    t_fit = float(probe.stdout.split()[3])
    nper = max(1, int(round(seconds_target / (10 * max(t_fit, 1e-3)))))
    t0 = time.perf_counter()
    procs = [subprocess.Popen([cli, "fits", input_path, str(2 + k * nper), str(1 + (k + 1) * nper), repr(conw_r)],
                              stdout=subprocess.PIPE, text=True) for k in range(cores)]
    nfits = 0
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            return None
        nfits += int(out.split()[1])
    wall = time.perf_counter() - t0
    return {"value": round(nfits / wall, 3), "unit": "fits/s", "cores": cores, "kind": kind,
            "sample": f"{label}: {cores} processes x {nper} trial period(s) x 10 poles of the same workload "
                      f"({nfits} fits, {wall:.1f} s wall), {os.path.relpath(cli, ROOT)}"}


def copy_calibration(mib):
    """One device-to-device copy of a known size: its FETCH_SIZE / WRITE_SIZE counter readings turn
    the fit kernel's readings into bytes."""
    import torch
    n = (mib << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 8, "fetch_kib_expected": n * 8 / 1024, "write_kib_expected": n * 8 / 1024}),
          flush=True)


def run_precision(ps, precision, windows, warmup, conw_r):
    """W untimed warmup steps (window 0), then one timed step per window."""
    import numpy as np
    ps.set_precision(precision)
    pole = None

    def step(n):
        nonlocal pole
        if pole is None or len(pole) != 10 * len(n):
            pole = 1 + np.arange(10 * len(n)) % 10
        return ps.fit_batch(ps.freq(np.repeat(n, 10)), pole, conw_r)

    for _ in range(warmup):
        step(windows[0])
    ps.kernel_time(reset=True)
    t0 = time.perf_counter()
    res = [step(n) for n in windows]
    wall = time.perf_counter() - t0
    kernel_ms, launches = ps.kernel_time()
    fits = 10 * sum(len(n) for n in windows)
    return {"fits": fits, "wall": wall, "value": fits / wall, "avg_launch_s": kernel_ms * 1e-3 / launches,
            "results": res}


def best_periods(res, n, ndata):
    """Per trial period the best of the 10 poles (lowest rms, first pole wins ties;
    period_search_BOINC.cpp:913-920): its rms and the line's chi2 = rms^2 * ndata."""
    import numpy as np
    dev = res["dev"].reshape(len(n), 10)
    best = dev[np.arange(len(n)), np.argmin(dev, axis=1)]
    return best, best * best * ndata


def agreement(r64, rmx, windows, ndata):
    import numpy as np
    same, errs, fit_errs = 0, [], []
    for a, b, n in zip(r64, rmx, windows):
        d64, c64 = best_periods(a, n, ndata)
        dmx, cmx = best_periods(b, n, ndata)
        same += int(np.argmin(d64) == np.argmin(dmx))
        errs.append(np.abs(cmx - c64) / c64)
        fit_errs.append(np.abs(b["chisq"] - a["chisq"]) / a["chisq"])
    errs, fit_errs = np.concatenate(errs), np.concatenate(fit_errs)
    return {"windows": len(windows), "same_best_period_windows": same,
            "period_chi2_rel_err_median": float(f"{np.median(errs):.3g}"),
            "period_chi2_rel_err_max": float(f"{np.max(errs):.3g}"),
            "fit_chi2_rel_err_median": float(f"{np.median(fit_errs):.3g}")}


def bench_diff(args):
    """The differentiation workload at the production shape: the 4096 parameter vectors of
    synth.config2_vectors, one ps_eval each -- the forward model with dyda of every parameter and the
    normal equations accumulated without storing J (the J-free kernels; the materialised-Jacobian
    kernels are timed beside them).  A step evaluates all 4096 vectors.  Accuracy against the reference
    core's values committed in tests/golden/diff_config2.npz (chi2, beta, alpha.u of all vectors; bright()
    rows of 16 vectors at 64 points) and against central finite differences of the GPU's own model."""
    import numpy as np

    from periodsearch_amd import PS_KERNEL_JFREE, PS_KERNEL_LEGACY, PeriodSearch, synth

    text, _ = synth.generate(synth.config1())
    ps = PeriodSearch.from_text(text).prepare(device=0)
    info = ps.info()
    fx = np.load(os.path.join(ROOT, "tests", "golden", "diff_config2.npz"))
    conw = float(fx["conw_r"])
    vecs = synth.config2_vectors(ps)
    rates = {}
    for name, form in (("legacy", PS_KERNEL_LEGACY), ("jfree", PS_KERNEL_JFREE)):
        ps.set_kernel(form)
        for _ in range(args.warmup):
            ps.eval(vecs, conw)
        ps.kernel_time(reset=True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            got = ps.eval(vecs, conw)
        wall = time.perf_counter() - t0
        kms, launches = ps.kernel_time()
        rates[name] = {"wall": wall, "avg_launch_s": kms * 1e-3 / launches, "got": got}
    r = rates["jfree"]
    got = r["got"]
    rel = lambda e, ref: float(np.max(e / np.abs(ref)))  # noqa: E731
    acc = {"chisq_vs_reference_max_rel": rel(np.abs(got["chisq"] - fx["chisq"]), fx["chisq"]),
           "beta_vs_reference_max_rel": float(np.max(np.abs(got["beta"] - fx["beta"]).max(1) /
                                                     np.abs(fx["beta"]).max(1))),
           "alpha_u_vs_reference_max_rel": float(np.max(np.abs(got["alpha"] @ fx["u"] - fx["alpha_u"]).max(1) /
                                                        np.abs(fx["alpha_u"]).max(1)))}
    raw = ps.eval(vecs[fx["bright_vec"]], conw, raw=True)
    d = raw["dyda"][:, fx["bright_pts"], :]
    acc["dyda_vs_reference_bright_max_rel"] = float(np.max(np.abs(d - fx["dyda"]).max((0, 1)) /
                                                           np.abs(fx["dyda"]).max((0, 1))))
    # central differences (relative step 1e-6; shape coefficients: of their basis scale; omega 1e-9 abs)
    ma, nc = info["ma"], info["ncoef"]
    span = np.abs(ps.geometry()["dsph"]).max(0)
    pts = np.array([0, 321, 1999, 2000, 2399])
    fd_err = 0.0
    for v in (5, 3000):
        a = vecs[v]
        an = ps.eval(a[None], conw, raw=True)["dyda"][0][pts]
        h = 1e-6 * np.maximum(np.abs(a), 1.0)
        h[:nc] = 1e-6 * np.maximum(np.abs(a[:nc]), 1.0 / span)
        h[nc + 2] = 1e-9
        plus, minus = np.tile(a, (ma, 1)), np.tile(a, (ma, 1))
        plus[np.arange(ma), np.arange(ma)] += h
        minus[np.arange(ma), np.arange(ma)] -= h
        fd = ((ps.eval(plus, conw, raw=True)["ymod"][:, pts] - ps.eval(minus, conw, raw=True)["ymod"][:, pts]) /
              (2 * h[:, None])).T
        fd_err = max(fd_err, float(np.max(np.abs(fd - an).max(0) / np.maximum(np.abs(an).max(0), 1e-300))))
    acc["dyda_vs_central_differences_max_rel"] = fd_err
    ps.close()
    f_eval = sum(flops_per_eval(info))
    nvec = len(vecs)
    achieved = nvec * f_eval / r["avg_launch_s"] / 1e12
    return {
        "metric": METRIC_CONFIG2, "value": round(nvec * args.steps / r["wall"], 2), "unit": "evaluations/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(r["wall"] / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded generator, periodsearch_amd/synth.py config1 + config2_vectors)",
        "config": {"workload": "config2_lmax6_288fac_ndata2400_4096vectors", "vectors_per_step": nvec,
                   "ndata": info["ndata"], "mfit": info["mfit"], "kernel": "jfree",
                   "legacy_evaluations_per_s": round(nvec * args.steps / rates["legacy"]["wall"], 2)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64, 4), "traffic": None, "flops_per_eval": f_eval,
                     "avg_launch_ms": round(r["avg_launch_s"] * 1e3, 3)},
        "accuracy": {k: float(f"{v:.3g}") for k, v in acc.items()},
        "cpu_baseline": None,
    }


def bench_batch(args):
    """The batched uneven workload: 512 seeded asteroids (synth.config4), a window of
    --batch-periods trial periods x 10 poles around each true period, all (asteroid, period, pole)
    fits as one ps_batch_run (a graph launch of persistent workgroups draining a device work queue).
    A step replays the same batch; the flops count each fit's own iteration number.  The same batch
    then runs in mixed precision (the mixed queue kernel), reported beside the FP64 value with its
    speedup and its agreement with the FP64 fits."""
    import numpy as np

    from periodsearch_amd import PS_FP64, PS_MIXED, Batch, PeriodSearch, synth

    searches, conw, prob, freq, pole, f_shape, f_ne = [], [], [], [], [], [], []
    for i in range(synth.CONFIG4_N):
        text, truth = synth.generate(synth.config4(i))
        ps = PeriodSearch.from_text(text).prepare(device=0)
        info = ps.info()
        n_true = int(round((info["freq_start"] - 1 / truth["period_h"]) / info["freq_step"])) + 1
        n = np.clip(n_true + np.arange(args.batch_periods) - args.batch_periods // 2, 1, info["nmax"])
        searches.append(ps)
        conw.append(info["conw_r0"])
        freq.extend(ps.freq(np.repeat(n, 10)))
        pole.extend(np.tile(np.arange(1, 11), len(n)))
        prob.extend([i] * (10 * len(n)))
        fs, fn = flops_per_eval(info)
        f_shape.append(fs)
        f_ne.append(fn)
    conw, prob, freq, pole, f_shape, f_ne = map(np.asarray, (conw, prob, freq, pole, f_shape, f_ne))

    def run(precision):
        for ps in searches:
            ps.set_precision(precision)
        with Batch(searches) as b:
            for _ in range(args.warmup):
                b.run(conw, prob, freq, pole)
            b.kernel_time(reset=True)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                res = b.run(conw, prob, freq, pole)
            wall = time.perf_counter() - t0
            kernel_ms, launches = b.kernel_time()
        return res, wall, kernel_ms * 1e-3 / launches

    res, wall, avg_launch = run(PS_FP64)
    mres, mwall, mavg = (None, None, None) if args.precision == "fp64" else run(PS_MIXED)
    for ps in searches:
        ps.close()
    nfits = len(freq)
    fs_tot = float(np.sum((res["niter"] + 1) * f_shape[prob]))
    fn_tot = float(np.sum((res["niter"] + 1) * f_ne[prob]))
    flops = fs_tot + fn_tot
    achieved = flops / avg_launch / 1e12
    line = {
        "metric": METRIC_CONFIG4, "value": round(nfits * args.steps / wall, 2), "unit": "fits/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded generator, periodsearch_amd/synth.py config4)",
        "config": {"workload": f"config4_{synth.CONFIG4_N}asteroids", "asteroids": synth.CONFIG4_N,
                   "fits_per_step": int(nfits), "periods_per_asteroid": args.batch_periods,
                   "lm_iterations_min_mean_max": [int(res["niter"].min()), round(float(res["niter"].mean()), 2),
                                                  int(res["niter"].max())],
                   "launch": "one graph: queue reset + persistent workgroups over a device work queue"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64, 4), "traffic": None, "flops_per_step": flops,
                     "avg_launch_ms": round(avg_launch * 1e3, 3)},
        "cpu_baseline": None,
    }
    if mres is not None:
        # flops from the mixed run's own iteration counts; the peak weighs the FP32 (shape contraction)
        # and FP64 (normal equations) parts by their flops
        mfs = float(np.sum((mres["niter"] + 1) * f_shape[prob]))
        mfn = float(np.sum((mres["niter"] + 1) * f_ne[prob]))
        mpeak = (mfs + mfn) / (mfs / PEAK_FP32 + mfn / PEAK_FP64)
        mach = (mfs + mfn) / mavg / 1e12
        ok = np.isfinite(res["chisq"]) & np.isfinite(mres["chisq"])
        err = np.abs(mres["chisq"][ok] - res["chisq"][ok]) / np.abs(res["chisq"][ok])
        # the best trial period of each asteroid's window (the lowest rms over its poles), FP64 vs mixed
        same = []
        for i in range(synth.CONFIG4_N):
            sel = np.flatnonzero(prob == i)
            per = np.round(freq[sel], 15)
            uniq = np.unique(per)
            if len(uniq) < 2:
                continue
            best = lambda r: uniq[np.argmin([np.nanmin(r["dev"][sel][per == u]) for u in uniq])]
            same.append(best(res) == best(mres))
        mval = nfits * args.steps / mwall
        line["mixed"] = {
            "value": round(mval, 2), "unit": "fits/s", "ms_per_step": round(mwall / args.steps * 1e3, 3),
            "speedup_vs_fp64": round(mval / line["value"], 3),
            "roofline": {"bound": "mfma", "achieved": round(mach, 3), "peak": round(mpeak, 2), "unit": "TFLOP/s",
                         "frac": round(mach / mpeak, 4), "traffic": None, "kernel": "legacy (mixed queue)",
                         "flops_per_step": mfs + mfn, "avg_launch_ms": round(mavg * 1e3, 3)},
            "agreement": {"asteroids": len(same), "same_best_period": int(np.sum(same)),
                          "fit_chi2_rel_err_median": float(f"{np.median(err):.3g}"),
                          "fit_chi2_rel_err_p90": float(f"{np.quantile(err, 0.9):.3g}"),
                          "same_iteration_count": round(float(np.mean(mres["niter"] == res["niter"])), 3)},
        }
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both",
                    help="both (default): FP64 as the value, mixed on the same fits beside it")
    ap.add_argument("--periods", type=int, default=256, help="trial periods per step (x 10 poles)")
    ap.add_argument("--workload", choices=["config1", "config2", "config3", "config4"], default="config1",
                    help="config1: dense photometry, Lmax 6 on 288 facets, 2400 points (the headline); "
                         "config2: 4096 model evaluations with dyda and the normal equations at the config1 shape; "
                         "config3: sparse survey, Lmax 6 on 512 facets, 5000 points; "
                         "config4: 512 asteroids of uneven size and shape as one batch (FP64)")
    ap.add_argument("--kernel", choices=["auto", "legacy", "jfree"], default="auto",
                    help="FP64 kernel form: auto (the autotuner picks, outside the timed region), or a fixed one")
    ap.add_argument("--batch-periods", type=int, default=2,
                    help="config4: trial periods per asteroid per step (x 10 poles)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU work per baseline process")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run one plain device copy of MIB MiB (the FETCH_SIZE/WRITE_SIZE scale factors)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); a scan over several GPUs runs one "
              "process per GPU on independent ranges of trial periods", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from periodsearch_amd import PS_FP64, PS_KERNEL_JFREE, PS_KERNEL_LEGACY, PS_MIXED, PeriodSearch, synth

    if args.workload == "config4":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.workload == "config2":
        print(json.dumps(bench_diff(args)), flush=True)
        return

    text, truth = synth.generate(getattr(synth, args.workload)())
    tmpdir = tempfile.mkdtemp(prefix="psbench_")
    input_path = os.path.join(tmpdir, "period_search_in")
    with open(input_path, "w") as f:
        f.write(text)
    ps = PeriodSearch.from_file(input_path).prepare(device=0)
    info = ps.info()
    conw_r = info["conw_r0"]
    # step k scans window k of the grid; window 0 is centred on the true period
    n_true = int(round((info["freq_start"] - 1 / truth["period_h"]) / info["freq_step"])) + 1
    n0 = max(1, n_true - args.periods // 2)
    windows = [1 + (n0 - 1 + k * args.periods + np.arange(args.periods)) % info["nmax"] for k in range(args.steps)]
    precisions = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    # the autotuner picks the FP64 kernel form on a sample of this problem's fits (outside the timed region)
    tune = None
    if "fp64" in precisions:
        ps.set_precision(PS_FP64)
        if args.kernel == "auto":
            kern, ms_legacy, ms_jfree = ps.autotune(conw_r)
            tune = {"kernel": "jfree" if kern == PS_KERNEL_JFREE else "legacy",
                    "ms_per_fit_legacy": round(ms_legacy, 5), "ms_per_fit_jfree": round(ms_jfree, 5) if ms_jfree else None}
        else:
            ps.set_kernel(PS_KERNEL_JFREE if args.kernel == "jfree" else PS_KERNEL_LEGACY)
            tune = {"kernel": args.kernel, "fixed": True}
    runs = {p: run_precision(ps, PS_FP64 if p == "fp64" else PS_MIXED, windows, args.warmup, conw_r)
            for p in precisions}
    ps.close()

    head = precisions[0]
    r = runs[head]
    niter = float(np.mean(np.concatenate([x["niter"] for x in r["results"]])))
    f_shape, f_ne = flops_per_eval(info)
    workload = f"{args.workload}_lmax6_{info['nfac']}fac_ndata{info['ndata']}"
    per_step = 10 * args.periods

    def roofline(p):
        x = runs[p]
        f_fit = (niter + 1) * (f_shape + f_ne)
        achieved = per_step * f_fit / x["avg_launch_s"] / 1e12
        if p == "fp64":
            peak = PEAK_FP64
        else:  # shape contraction on FP32 MFMA, normal equations on FP64 MFMA: the time-weighted peak
            peak = (f_shape + f_ne) / (f_shape / PEAK_FP32 + f_ne / PEAK_FP64)
        form = tune["kernel"] if (p == "fp64" and tune) else "legacy"
        return {"bound": "mfma", "achieved": round(achieved, 3), "peak": round(peak, 2), "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": traffic_per_launch(workload, p, form), "kernel": form,
                "flops_per_fit": f_fit, "avg_launch_ms": round(x["avg_launch_s"] * 1e3, 3)}

    line = {
        "metric": METRIC if args.workload == "config1" else METRIC_CONFIG3,
        "value": round(r["value"], 2),
        "unit": "fits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["wall"] / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32+f64",
        "data": f"synthetic (seeded generator, periodsearch_amd/synth.py {args.workload})",
        "config": {"workload": workload, "ndata": info["ndata"], "lmax": 6, "nfac": info["nfac"],
                   "mfit": info["mfit"], "lm_iterations": niter, "periods_per_step": args.periods,
                   "fits_per_step": per_step, "precision": head, "first_period_index": int(windows[0][0]),
                   "trial_periods_per_s": round(r["value"] / 10, 2)},
        "roofline": roofline(head),
        "autotune": tune,
    }
    if len(precisions) == 2:
        m = runs["mixed"]
        line["mixed"] = {"value": round(m["value"], 2), "unit": "fits/s",
                         "ms_per_step": round(m["wall"] / args.steps * 1e3, 3),
                         "speedup_vs_fp64": round(m["value"] / r["value"], 3), "roofline": roofline("mixed"),
                         "agreement": agreement(r["results"], m["results"], windows, info["ndata"])}
    if not args.no_cpu_baseline:
        fma = cpu_baseline(REF_FMA_CLI, "reference AVX-FMA build", "reference", input_path, conw_r,
                           args.cpu_seconds)
        plain = cpu_baseline(REF_CLI, "reference plain-C build", "reference", input_path, conw_r, args.cpu_seconds)
        if plain is None:
            plain = cpu_baseline(ORACLE_CLI, "FP64 restatement of the plain-C core", "port", input_path, conw_r,
                                 args.cpu_seconds)
        line["cpu_baseline"] = fma or plain
        line["cpu_baseline_plain"] = plain if fma else None
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
