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

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, 8 TB/s
FP64_PEAK_TFS = 78.6    # MI355X FP64 vector
PATH_BYTES_PER_VOXEL_ITER = 768   # FP64: FFT 96 + Green/inverse 96 + Newton reads 432 + writes 144


PROF_EVERY = 10


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
    if name == "nr":                         # read E, S, EP, CRSS, R, crystal-frame K_c (deviatoric
        # triangle), ORI, PH; write S, P
        return (6 + 6 + 5 + 12 * (2 if two_crss else 1) + 9 + 15) * 8 + 4 + 1 + (6 + 6) * 8
    return 0.0


def make_case(d, steps):
    from voce_amd import inputs
    c = inputs.config_case(2)
    c.n_steps = steps
    log("generating 64^3 / 125-grain Voronoi mesh in", d)
    inputs.write_case(d, c)
    return os.path.join(d, "input"), c


def run_gpu(path, base, precision, steps, warmup, prof=True, tune=True):
    import voce_amd
    cs = voce_amd.load_case(path, base)
    if warmup > 0:
        w = voce_amd.Solver.from_case(cs, precision)
        for i in range(warmup):
            w.increment()
        w.close()
        log(precision, "warmup", warmup, "increments done")
    s = voce_amd.Solver.from_case(cs, precision)
    tuned = s.autotune() if tune else None
    if tuned:
        log(precision, "tuned FFT launch choice", tuned)
    if prof:
        # events on every PROF_EVERY-th FFT iteration: each recorded event
        # costs queue time, a sparse sample keeps the timed run honest
        s.prof_enable(True, every=PROF_EVERY)
    reps = []
    t_inc = []
    t0 = time.perf_counter()
    for k in range(steps):
        t1 = time.perf_counter()
        reps.append(s.increment())
        t_inc.append(time.perf_counter() - t1)
        if (k + 1) % 20 == 0:
            log(precision, "increment", k + 1, "iters", reps[-1]["iter"], "elapsed %.2f s" % (time.perf_counter() - t0))
    wall = time.perf_counter() - t0
    kp = s.prof_read() if prof else {}
    V = s.nvox
    s.close()
    run_gpu.tuned = tuned
    run_gpu.t_inc = t_inc
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


def run_plane_strain(d, steps, warmup, precisions=("mixed", "fp64"), n=128):
    """128^3 / 1000-grain plane-strain compression (velocity gradient
    diag(1, 0, -1) 1e-3/s, dt 1 s): MIXED, and FP64 on the same seed for the
    error of the macroscopic stress curve and of the final stress field;
    n^3 voxels instead of 128^3 with n."""
    import voce_amd
    from voce_amd import inputs
    c = inputs.config_case(3)
    c.dims = (n, n, n)
    c.n_steps = steps + warmup
    d3 = os.path.join(d, "plane_strain_%d" % n)
    os.makedirs(d3, exist_ok=True)
    log("generating %d^3 / 1000-grain Voronoi mesh in" % n, d3)
    inputs.write_case(d3, c)
    cs = voce_amd.load_case(os.path.join(d3, "input"), d3)
    res, curves, fields = {}, {}, {}
    for prec in precisions:
        s = voce_amd.Solver.from_case(cs, prec)
        for _ in range(warmup):
            s.increment()
        s.close()
        # a fresh context, tuned (outside the clock), timed
        s = voce_amd.Solver.from_case(cs, prec)
        tuned = s.autotune()
        reps = []
        t0 = time.perf_counter()
        for _ in range(steps):
            reps.append(s.increment())
        wall = time.perf_counter() - t0
        it = sum(r["iter"] for r in reps)
        fields[prec] = s.field("sig")
        curves[prec] = reps
        V = s.nvox
        s.close()
        res[prec] = {"voxel_iterations_per_s": round(V * it / wall, 1), "fft_iterations": it,
                     "ms_per_increment": round(wall / steps * 1e3, 3), "wall_s": round(wall, 3),
                     "fft_launch_choice": tuned}
        log("%d^3 %s: %d increments, %d iterations, %.3f s, %.4g voxel-it/s" % (n, prec, steps, it, wall, V * it / wall))
    cs.free()
    out = {"workload": "single RVE, %d^3 voxels, 1000 Voronoi grains (seed 3), FCC Cu-like, plane-strain "
                        "compression L = diag(1, 0, -1) 1e-3/s, dt 1 s, Err 1e-5, IterMax 100; timed: increments "
                        "1-%d (%.1f %% strain) of a fresh context after %d warm-up increments on another"
                        % (n, steps, steps * 0.1, warmup),
           "voxels": V, "increments": steps}
    out.update(res)
    if len(res) == 2:
        out["mixed_over_fp64"] = round(res["mixed"]["voxel_iterations_per_s"] / res["fp64"]["voxel_iterations_per_s"], 3)
        out["macro_stress_max_rel_err_vs_fp64"] = max(
            float(np.abs(a["sig_avg"] - b["sig_avg"]).max() / np.abs(b["sig_avg"]).max())
            for a, b in zip(curves["mixed"], curves["fp64"]))
        fm, ff = fields["mixed"], fields["fp64"]
        out["sigma_field_max_abs_err_over_max_vs_fp64"] = float(np.abs(fm - ff).max() / np.abs(ff).max())
    return out


def run_batched(d, precision, n_rve, steps, log_every=5):
    """Config-4 style workload: n_rve RVEs of uneven size, each with its own
    velocity gradient, advanced `steps` increments together."""
    import voce_amd
    from voce_amd import inputs
    specs = inputs.batch_specs(n=n_rve)
    t0 = time.perf_counter()
    cases = inputs.load_specs(os.path.join(d, "batch_%s" % precision), specs,
                              inputs.full_velocity_case(time_step=1.0, err=1e-5, iter_max=100))
    b = voce_amd.Batch(cases, precision, tune=True)
    t_setup = time.perf_counter() - t0
    for c in cases:
        c.free()
    V = [sp.dims[0] * sp.dims[1] * sp.dims[2] for sp in specs]
    log("batched %s: %d RVEs, %d voxels, setup %.1f s" % (precision, n_rve, sum(V), t_setup))
    iters = 0
    reps_all = []
    t0 = time.perf_counter()
    for k in range(steps):
        reps = b.increment()
        reps_all.append(reps)
        iters += sum(v * r["iter"] for v, r in zip(V, reps))
        if (k + 1) % log_every == 0:
            log("batched %s increment %d, max iters %d, elapsed %.2f s" % (precision, k + 1,
                                                                           max(r["iter"] for r in reps),
                                                                           time.perf_counter() - t0))
    wall = time.perf_counter() - t0
    info = b.info()
    b.close()
    return {"rve_increments_per_s": round(n_rve * steps / wall, 2), "voxel_iterations_per_s": round(iters / wall, 1),
            "wall_s": round(wall, 3), "setup_s": round(t_setup, 2), "rves": n_rve, "increments": steps,
            "voxels": sum(V), "mean_iterations": round(float(np.mean([r["iter"] for rr in reps_all for r in rr])), 2),
            "graph_replays": info["graph_replays"], "device_bytes": info["device_bytes"]}, reps_all


def run_served(d, n_ids, n_req, rates, audit_pct=1.0, seed=5, precision="mixed"):
    """Config-5 style stream: n_ids resident RVEs, Poisson arrivals of
    one-increment requests (random traceless L, |L| = 1e-3/s, dt 1 s),
    served in `precision` with an FP64 audit of audit_pct % of the requests.
    The first stream (all requests at t = 0) measures capacity; the same
    seed gives the same ids and requests in either precision."""
    import voce_amd
    from voce_amd import inputs
    specs = inputs.serve_specs(n_ids=n_ids)
    base = inputs.full_velocity_case(time_step=1.0, err=1e-5, iter_max=100)
    t0 = time.perf_counter()
    srv = voce_amd.Server(n_ids, precision, tune=True)
    inputs.write_case(os.path.join(d, "serve"), base, grain=np.ones(base.nvox, dtype=np.int32),
                      euler=np.zeros((1, 3)))
    path = os.path.join(d, "serve", "input")
    V = {}
    for i, sp in enumerate(specs):
        cs = voce_amd.load_case_voronoi(path, os.path.join(d, "serve"), sp.dims, sp.n_grains, sp.seed)
        cs.set_velgrad_full(np.zeros((3, 3)))
        srv.add_rve(i, cs)
        cs.free()
        V[i] = sp.dims[0] * sp.dims[1] * sp.dims[2]
        if (i + 1) % 256 == 0:
            log("served: registered %d ids (%.1f s)" % (i + 1, time.perf_counter() - t0))
    t_setup = time.perf_counter() - t0
    info0 = srv.info()
    log("served %s: %d ids resident, %.1f GB, setup %.1f s" % (precision, n_ids, info0["resident_bytes"] / 1e9,
                                                                t_setup))
    rng = inputs.SplitMix64(seed)

    def stream(rate, n, tag0):
        """n requests; rate None = all at t = 0 (capacity)."""
        arr, reqs = [], []
        t = 0.0
        for i in range(n):
            if rate:
                t += -np.log(1.0 - rng.uniform()) / rate
            arr.append(t)
            rid = min(n_ids - 1, int(rng.uniform() * n_ids))
            aud = (((tag0 + i) * 0x9E3779B1) % 10007) < audit_pct / 100.0 * 10007
            reqs.append((rid, inputs.request_velgrad(rng), aud))
        res = {}
        t0 = time.perf_counter()
        i = 0
        last = t0
        while len(res) < n:
            now = time.perf_counter()
            while i < n and t0 + arr[i] <= now:
                rid, L, aud = reqs[i]
                srv.submit(rid, L, aud, tag0 + i)
                i += 1
            inflight = srv.step()
            for r in srv.poll():
                res[r["tag"] - tag0] = r
            if inflight == 0 and srv.info()["queued"] == 0 and i < n:
                wait = t0 + arr[i] - time.perf_counter()
                if wait > 0:
                    time.sleep(min(wait, 0.01))
            if time.perf_counter() - last > 20:
                last = time.perf_counter()
                log("served: %d/%d done" % (len(res), n))
        wall = time.perf_counter() - t0
        lat = np.array([res[k]["t_done"] - (t0 + arr[k]) for k in range(n)]) * 1e3
        vi = sum(V[reqs[k][0]] * res[k]["iter"] for k in range(n))
        aud = [res[k]["audit_sig_rel"] for k in range(n) if res[k]["audited"]]
        return {"offered_per_s": (round(rate, 2) if rate else None), "requests": n,
                "increments_per_s": round(n / wall, 2), "voxel_iterations_per_s": round(vi / wall, 1),
                "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
                "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
                "mean_iterations": round(float(np.mean([res[k]["iter"] for k in range(n)])), 2),
                "audits": len(aud), "audit_sig_rel_max": (max(aud) if aud else None)}

    cap = stream(None, n_req, 0)
    log("served %s capacity: %.1f increments/s, p50 %.1f ms p99 %.1f ms" % (precision, cap["increments_per_s"],
                                                                        cap["latency_ms_p50"], cap["latency_ms_p99"]))
    sweep = []
    for j, f in enumerate(rates):
        r = stream(f * cap["increments_per_s"], n_req, (j + 1) * 10_000_000)
        r["load"] = f
        sweep.append(r)
        log("served at %.2fx capacity: %.1f inc/s, p50 %.1f ms, p99 %.1f ms" % (f, r["increments_per_s"],
                                                                                r["latency_ms_p50"],
                                                                                r["latency_ms_p99"]))
    info = srv.info()
    srv.close()
    return {"ids": n_ids, "precision": precision, "resident_gb": round(info0["resident_bytes"] / 1e9, 2),
            "setup_s": round(t_setup, 1), "capacity": cap, "sweep": sweep, "steps": info["steps"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=10)
    ap.add_argument("--batch-rves", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--serve-ids", type=int, default=4096)
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-serve-fp64", action="store_true")
    ap.add_argument("--ps128-steps", type=int, default=100)
    ap.add_argument("--no-ps128", action="store_true")
    # only the 128^3 plane-strain leg, in the named precisions (profiling);
    # --ps-n sets its grid (n^3) there
    ap.add_argument("--ps128-only", choices=("mixed", "fp64", "both"))
    ap.add_argument("--ps-n", type=int, default=128)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the RVE lives whole on one MI355X and nothing here "
              "spans GPUs (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import voce_amd
    voce_amd.load_library()
    d = tempfile.mkdtemp(prefix="voce_bench_")
    if args.ps128_only:
        precs = ("mixed", "fp64") if args.ps128_only == "both" else (args.ps128_only,)
        print(json.dumps({"plane_strain_%d" % args.ps_n: run_plane_strain(d, args.ps128_steps, 1, precs, args.ps_n)}),
              flush=True)
        return
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
    # measured HBM bytes per launch of that kernel: rocprofv3 FETCH_SIZE/WRITE_SIZE
    # (scaled by a copy-kernel calibration) committed with the profile summary
    traffic, tsrc, flop = None, None, None
    tf = os.path.join(ROOT, "profiles", "iteration_kernels_64cube_fp64.traffic.json")
    if os.path.exists(tf):
        try:
            tj = json.load(open(tf))
            traffic = tj["bytes_per_launch"].get(dom)
            flop = tj.get("fp64_flop_per_launch", {}).get(dom)
            tsrc = "profiles/iteration_kernels_64cube_fp64.md"
        except Exception:
            traffic = None
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": (round(traffic) if traffic else None),
            "traffic_source": tsrc, "algorithmic_bytes": round(bpv * V),
            "bytes_per_voxel": round(bpv, 2), "avg_launch_ms": round(avg_ms, 5)}
    # The Newton kernel is FP64-VALU bound rather than HBM bound: its FP64
    # work per launch (rocprofv3 SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64, 64
    # lanes per wave instruction, FMA = 2) over the live average duration,
    # against the MI355X FP64 vector peak
    roof_fp64 = None
    if flop and avg_ms > 0:
        tf_s = flop / (avg_ms * 1e-3) / 1e12
        roof_fp64 = {"bound": "fp64-valu", "kernel": dom, "achieved": round(tf_s, 2), "peak": FP64_PEAK_TFS,
                     "unit": "TFLOP/s", "frac": round(tf_s / FP64_PEAK_TFS, 4), "flop_per_launch": round(flop),
                     "flop_source": tsrc}
    # whole FFT iteration at the per-voxel-iteration byte count of the
    # reference path's single-pass design (768 B FP64, every operand once)
    path_gbs = vox_it * PATH_BYTES_PER_VOXEL_ITER / 1e9
    roof_path = {"bytes_per_voxel_iteration": PATH_BYTES_PER_VOXEL_ITER, "achieved": round(path_gbs, 1),
                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(path_gbs / HBM_PEAK_GBS, 4)}
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
                               "1e-3/s, dt 0.5, Err 1e-5, IterMax 100 (the 100-increment run); timed: increments "
                               "1-%d of a fresh context after %d warm-up increments on another" % (args.steps,
                                                                                                   args.warmup),
                   "voxels": V, "increments": args.steps, "precision": "fp64"},
        "rve_increments_per_s": round(args.steps / wall, 3),
        "fft_iterations": iters,
        "increment_ms": {"p50": round(float(np.percentile(run_gpu.t_inc, 50)) * 1e3, 3),
                         "p99": round(float(np.percentile(run_gpu.t_inc, 99)) * 1e3, 3),
                         "what": "one increment of the single 64^3 FP64 RVE, call to return"},
        "roofline": roof,
        "roofline_fp64": roof_fp64,
        "roofline_path": roof_path,
        "kernels": per_kernel,
        "fft_launch_choice": run_gpu.tuned,
    }

    if not args.no_mixed:
        mreps, mwall, _, _ = run_gpu(path, d, "mixed", args.steps, 0, prof=False)
        miters = sum(r["iter"] for r in mreps)
        err = max(float(np.abs(a["sig_avg"] - b["sig_avg"]).max() / np.abs(b["sig_avg"]).max())
                  for a, b in zip(mreps, reps))
        out["mixed"] = {"value": round(V * miters / mwall, 1), "unit": "voxel-iterations/s",
                        "over_fp64": round((V * miters / mwall) / vox_it, 3),
                        "rve_increments_per_s": round(args.steps / mwall, 3), "fft_iterations": miters,
                        "ms_per_step": round(mwall / args.steps * 1e3, 3),
                        "macro_stress_max_rel_err_vs_fp64": err}
        log("mixed: %.3f s, %.4g voxel-it/s, macro rel err %.2e" % (mwall, V * miters / mwall, err))

    if not args.no_ps128:
        out["plane_strain_128"] = run_plane_strain(d, args.ps128_steps, 1)

    rve_inc_fp64 = out["rve_increments_per_s"]
    if not args.no_batch:
        bf, reps64 = run_batched(d, "fp64", args.batch_rves, args.batch_steps)
        bm, repsm = run_batched(d, "mixed", args.batch_rves, args.batch_steps)
        err = 0.0
        for a, b in zip(repsm, reps64):
            for x, y in zip(a, b):
                err = max(err, float(np.abs(x["sig_avg"] - y["sig_avg"]).max() / np.abs(y["sig_avg"]).max()))
        bm["macro_stress_max_rel_err_vs_fp64"] = err
        out["batched"] = {"workload": "%d RVEs, N^3 with N in {16,32,48,64}, 8-125 grains, own random L = D + W at "
                                      "1e-3/s, dt 1 s, %d increments, Err 1e-5, IterMax 100" % (args.batch_rves,
                                                                                                  args.batch_steps),
                          "fp64": bf, "mixed": bm,
                          # same workload, precision alone
                          "mixed_over_fp64_voxel_iterations": round(bm["voxel_iterations_per_s"] /
                                                                    bf["voxel_iterations_per_s"], 3),
                          # against the single 64^3 FP64 RVE: RVE-increments (a mix of smaller RVEs) and
                          # voxel-iterations (work done)
                          "mixed_rve_inc_over_plain_fp64_single": round(bm["rve_increments_per_s"] / rve_inc_fp64, 2),
                          "mixed_voxel_it_over_plain_fp64_single": round(bm["voxel_iterations_per_s"] / vox_it, 3)}
    if not args.no_serve:
        sv = run_served(d, args.serve_ids, args.serve_requests, (0.25, 0.5, 1.0, 2.0))
        sv["workload"] = ("%d resident ids, N^3 with N in {16,32,64}, 8-64 grains; Poisson arrivals of one-increment "
                          "requests, random traceless L (|L| = 1e-3/s), dt 1 s; mixed precision, FP64 audit of 1%%"
                          % args.serve_ids)
        sv["capacity_over_plain_fp64_single"] = round(sv["capacity"]["increments_per_s"] / rve_inc_fp64, 2)
        sv["capacity_voxel_it_over_plain_fp64_single"] = round(sv["capacity"]["voxel_iterations_per_s"] / vox_it, 3)
        if not args.no_serve_fp64:
            # the same ids and the same capacity stream served in FP64
            sf = run_served(d, args.serve_ids, args.serve_requests, (), audit_pct=0.0, precision="fp64")
            sv["fp64_capacity"] = sf["capacity"]
            sv["fp64_setup_s"] = sf["setup_s"]
            sv["mixed_over_fp64_capacity"] = round(sv["capacity"]["increments_per_s"] /
                                                   sf["capacity"]["increments_per_s"], 3)
        out["served"] = sv

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
