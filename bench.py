#!/usr/bin/env python3
"""Benchmark of the period scan hot path: trial-period Levenberg-Marquardt fits per second at
Lmax = Mmax = 6 on 288 facets (the production shape: 40 dense lightcurves x 50 points + 400 sparse
calibrated points, ndata = 2400, 55 fitted parameters, stop condition 50 = fixed 50 LM iterations).

A step is one ps_fit_batch call over a window of trial periods x 10 initial poles, inputs resident
in HBM (uploaded once by ps_prepare); ps_fit_batch returns after the library's stream has
synchronised, so the wall clock around the K steps covers all device work.  The roofline figure uses
the fit kernel's own duration, measured with HIP events recorded on the stream it is launched on.

    python bench.py [--steps K] [--warmup W] [--precision fp64|mixed] [--periods P] [--workload config1|config3]

The default workload (config1) is the headline; config3 (a sparse survey fitted on 512 facets, Dsph
read through L2) is a second shape, reported under its own metric name.
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
PEAK_TFLOPS = {"fp64": 78.6, "mixed": 157.3}  # MI355X dense FP64 / FP32 matrix peaks
ORACLE_CLI = os.path.join(ROOT, "oracle", "ps_oracle_cli")
REF_CLI = os.path.join(ROOT, "oracle", "_ref", "ps_ref_cli")


def algorithmic_flops_per_eval(info):
    """Shape-Jacobian contraction + J^T W J lower triangle with diagonal + J^T W r, per model
    evaluation: N * (2 Nf Nc + mfit (mfit + 1) + 2 mfit), N = ndata + 3 convexity rows."""
    n, nf, nc, mf = info["npts"], info["nfac"], info["ncoef"], info["mfit"]
    return n * (2 * nf * nc + mf * (mf + 1) + 2 * mf)


def traffic_per_launch(workload, precision):
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t.get("hbm_bytes_per_launch")


def cpu_baseline(input_path, conw_r, seconds_target=20.0):
    """The reference's own plain-C core (oracle/_ref/ps_ref_cli, compiled from the reference sources by
    oracle/Makefile; kind "reference") or, where that was not built, the oracle's FP64 restatement
    (oracle/ps_oracle_cli; kind "port"), single-threaded as the reference is, one process per host core
    (at most 16) on disjoint trial periods of the same workload -- the reference's own data-parallel
    split.  Each process runs whole periods (10 poles) for about seconds_target."""
    if os.path.exists(REF_CLI):
        cli, kind = REF_CLI, "reference"
    elif os.path.exists(ORACLE_CLI):
        cli, kind = ORACLE_CLI, "port"
    else:
        return None
    cores = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()))
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
            "sample": f"{cores} processes x {nper} trial period(s) x 10 poles of the same workload "
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--compare", action="store_true",
                    help="also run the other precision on the same fits: speedup and best-period agreement")
    ap.add_argument("--periods", type=int, default=256, help="trial periods per step (x 10 poles)")
    ap.add_argument("--workload", choices=["config1", "config3"], default="config1",
                    help="config1: dense photometry, Lmax 6 on 288 facets, 2400 points (the headline); "
                         "config3: sparse survey, Lmax 6 on 512 facets, 5000 points")
    ap.add_argument("--no-cpu-baseline", action="store_true")
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

    from periodsearch_amd import PS_FP64, PS_MIXED, PeriodSearch, synth

    text, _ = synth.generate(getattr(synth, args.workload)())
    tmpdir = tempfile.mkdtemp(prefix="psbench_")
    input_path = os.path.join(tmpdir, "period_search_in")
    with open(input_path, "w") as f:
        f.write(text)
    ps = PeriodSearch.from_file(input_path).prepare(device=0,
                                                    precision=PS_MIXED if args.precision == "mixed" else PS_FP64)
    info = ps.info()
    conw_r = info["conw_r0"]
    per_step = args.periods * 10

    def step(k):
        n = 1 + (k * args.periods + np.arange(per_step) // 10) % info["nmax"]
        pole = 1 + np.arange(per_step) % 10
        return ps.fit_batch(ps.freq(n), pole, conw_r)

    for k in range(args.warmup):
        step(k)
    ps.kernel_time(reset=True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        res = step(args.warmup + k)
    wall = time.perf_counter() - t0
    kernel_ms, launches = ps.kernel_time()
    fits = per_step * args.steps
    value = fits / wall

    niter = float(np.mean(res["niter"]))
    f_fit = (niter + 1) * algorithmic_flops_per_eval(info)
    avg_launch_s = kernel_ms * 1e-3 / launches
    achieved = per_step * f_fit / avg_launch_s / 1e12
    peak = PEAK_TFLOPS[args.precision]
    workload = f"{args.workload}_lmax6_{info['nfac']}fac_ndata{info['ndata']}"
    line = {
        "metric": METRIC if args.workload == "config1" else METRIC_CONFIG3,
        "value": round(value, 2),
        "unit": "fits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": f"synthetic (seeded generator, periodsearch_amd/synth.py {args.workload})",
        "config": {"workload": workload, "ndata": info["ndata"], "lmax": 6, "nfac": info["nfac"],
                   "mfit": info["mfit"], "lm_iterations": niter, "periods_per_step": args.periods,
                   "fits_per_step": per_step, "precision": args.precision,
                   "trial_periods_per_s": round(value / 10, 2)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4),
                     "traffic": traffic_per_launch(workload, args.precision),
                     "flops_per_fit": f_fit, "avg_launch_ms": round(avg_launch_s * 1e3, 3)},
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(input_path, conw_r),
    }
    ps.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
