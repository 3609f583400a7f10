#!/usr/bin/env python3
"""Benchmark of the EVP-FFT hot path on one MI355X.

Workload (the single-RVE case the metric is quoted on): a 64^3 voxel grid
of 125 seeded periodic-Voronoi grains of Cu-like FCC crystal (C11 168.4,
C12 121.4, C44 75.4 GPa; 12 {111}<110> systems, n = 20, gamma0 = 1/s; Voce
tau0 16, tau1 2, theta0 250, theta1 50 MPa), uniaxial tension along x with
stress-free lateral faces, 100 increments of dt = 0.5 s at 1e-3/s (5 %),
Err 1e-5, IterMax 100.  Inputs are generated synthetically (data:
"synthetic") in the reference's own file formats.

One "step" is one strain increment of that RVE: every FFT fixed-point
iteration to convergence plus the close-out.  --warmup W increments run on
a throw-away context; then a fresh context times --steps K increments
(setup excluded, inputs already resident in HBM).  The C ABI returns from
voce_increment only after its stream has drained, so the wall clock around
the K calls is the device time plus host orchestration.

metric value = voxel-iterations/s (FP64) = V * sum(iterations) / wall.
Also reported: RVE-increments/s, the same run in MIXED precision and its
macro-stress error against FP64, per-kernel averages from HIP events
recorded on the solver's stream during the timed region, the roofline of
the dominant kernel and a CPU baseline (the oracle on the host cores, a
bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, 8 TB/s


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def kernel_bytes_per_voxel(name, nx, ny, nz, f32, two_crss):
    """Algorithmic HBM bytes per voxel of one launch of each iteration kernel:
    every operand streamed once (layout in voce_amd/csrc/voce_internal.h)."""
    c = 8 if f32 else 16                     # complex FFT element
    spec = (nx * ny * (nz // 2 + 1)) / float(nx * ny * nz)   # half-spectrum points per voxel
    if name == "plane_fwd":                  # read sigma (6 f64), write 6 half spectra
        return 6 * 8 + 6 * c * spec
    if name == "xgreen":                     # read + write 6 spectra, read + write 3 u sums
        return 2 * 6 * c * spec + 2 * 3 * c * spec
    if name == "plane_inv":                  # read 6 spectra, read + write E (6 f64)
        return 6 * c * spec + 2 * 6 * 8
    if name == "nr":                         # read E, S, EP, CRSS, R, ORI, PH; write S, P
        return (6 + 6 + 5 + 12 * (2 if two_crss else 1) + 9) * 8 + 4 + 1 + (6 + 6) * 8
    return 0.0


def make_case(d, steps):
    from voce_amd import inputs
    c = inputs.config_case(2)
    c.n_steps = steps
    log("generating 64^3 / 125-grain Voronoi mesh in", d)
    inputs.write_case(d, c)
    return os.path.join(d, "input"), c


def run_gpu(path, base, precision, steps, warmup, prof=True):
    import voce_amd
    cs = voce_amd.load_case(path, base)
    if warmup > 0:
        w = voce_amd.Solver.from_case(cs, precision)
        for i in range(warmup):
            w.increment()
        w.close()
        log(precision, "warmup", warmup, "increments done")
    s = voce_amd.Solver.from_case(cs, precision)
    if prof:
        s.prof_enable(True)
    reps = []
    t0 = time.perf_counter()
    for k in range(steps):
        reps.append(s.increment())
        if (k + 1) % 20 == 0:
            log(precision, "increment", k + 1, "iters", reps[-1]["iter"], "elapsed %.2f s" % (time.perf_counter() - t0))
    wall = time.perf_counter() - t0
    kp = s.prof_read() if prof else {}
    V = s.nvox
    s.close()
    return reps, wall, kp, V


def cpu_baseline(path, base, sample_iters, threads):
    """The oracle (FP64 C restatement of the reference) on the host cores:
    the first increment of the same case truncated to `sample_iters` FFT
    iterations."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    txt = open(path).read().splitlines()
    txt = [("IterMax\t%d" % sample_iters) if ln.startswith("IterMax") else ln for ln in txt]
    p2 = path + ".cpu"
    open(p2, "w").write("\n".join(txt) + "\n")
    j = po.Job(p2, base, threads=threads)
    t0 = time.perf_counter()
    r = j.increment()
    dt = time.perf_counter() - t0
    V = j.nvox
    j.close()
    return V * r["iter"] / dt, dt, r["iter"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=10)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the RVE lives whole on one MI355X and nothing here "
              "spans GPUs (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import voce_amd
    voce_amd.load_library()
    d = tempfile.mkdtemp(prefix="voce_bench_")
    path, case = make_case(d, args.steps)
    nx, ny, nz = case.dims

    reps, wall, kp, V = run_gpu(path, d, "fp64", args.steps, args.warmup)
    iters = sum(r["iter"] for r in reps)
    vox_it = V * iters / wall
    log("fp64: %d increments, %d iterations, %.3f s, %.4g voxel-it/s" % (args.steps, iters, wall, vox_it))

    # dominant kernel and its roofline
    kern = {k: (ms / n if n else 0.0, n, ms) for k, (ms, n) in kp.items()}
    dom = max(kern, key=lambda k: kern[k][2])
    avg_ms = kern[dom][0]
    bpv = kernel_bytes_per_voxel(dom, nx, ny, nz, False, case.material.tau0f != case.material.tau0b)
    achieved = bpv * V / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_per_voxel": round(bpv, 2), "avg_launch_ms": round(avg_ms, 5)}
    per_kernel = {k: {"avg_ms": round(v[0], 5), "launches": v[1], "share": round(v[2] / max(1e-12, sum(x[2] for x in kern.values())), 4)}
                  for k, v in kern.items()}

    out = {
        "metric": "voxel-iterations/s + RVE-increments/s (p50/p99 latency), mixed vs FP64, 1 GPU",
        "value": round(vox_it, 1),
        "unit": "voxel-iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "single RVE, 64^3 voxels, 125 Voronoi grains (seed 2), FCC Cu-like, x-tension "
                               "1e-3/s, 100 increments dt 0.5, Err 1e-5, IterMax 100",
                   "voxels": V, "increments": args.steps, "precision": "fp64"},
        "rve_increments_per_s": round(args.steps / wall, 3),
        "fft_iterations": iters,
        "increment_ms_p50": None,
        "roofline": roof,
        "kernels": per_kernel,
    }

    if not args.no_mixed:
        mreps, mwall, _, _ = run_gpu(path, d, "mixed", args.steps, 0, prof=False)
        miters = sum(r["iter"] for r in mreps)
        import numpy as np
        err = max(float(np.abs(a["sig_avg"] - b["sig_avg"]).max() / np.abs(b["sig_avg"]).max())
                  for a, b in zip(mreps, reps))
        out["mixed"] = {"value": round(V * miters / mwall, 1), "unit": "voxel-iterations/s",
                        "rve_increments_per_s": round(args.steps / mwall, 3), "fft_iterations": miters,
                        "ms_per_step": round(mwall / args.steps * 1e3, 3),
                        "macro_stress_max_rel_err_vs_fp64": err}
        log("mixed: %.3f s, %.4g voxel-it/s, macro rel err %.2e" % (mwall, V * miters / mwall, err))

    if not args.no_cpu:
        try:
            ncpu = len(os.sched_getaffinity(0))
        except Exception:
            ncpu = os.cpu_count() or 1
        threads = max(1, min(16, ncpu))
        log("cpu baseline: oracle, %d threads, %d iterations of increment 1" % (threads, args.cpu_iters))
        v, t, n = cpu_baseline(path, d, args.cpu_iters, threads)
        out["cpu_baseline"] = {"value": round(v, 1), "unit": "voxel-iterations/s", "cores": threads, "kind": "port",
                               "sample": "oracle (FP64 C restatement of the reference), increment 1 of the same 64^3 "
                                         "case truncated to %d FFT iterations, %.1f s" % (n, t)}
        out["gpu_over_cpu"] = round(vox_it / v, 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
