"""Writes the oracle's FP64 stress-strain curves that the GPU tests compare
against for the two long cases (too long to recompute on the CPU in every
GPU test run):

  config1_fp64_50incr.npz   16^3, 8 grains (seed 1), x tension with stress-free
                            lateral faces, dt 1, all 50 increments
  config2_fp64_4incr.npz    64^3, 125 grains (seed 2), x tension, dt 0.5, the
                            first 4 increments (hardening from the third)

Per increment: iteration count, running total, the six macro tensors, the
vm.out scalars and the per-iteration Err_e, Err_s, s_vm rows.  At the last
increment: sig, crss and tranmat at 1024 voxels on a fixed stride, and the
per-grain means of sig and crss.  The inputs are the repository's own
seeded Voronoi cases (voce_amd.inputs.config_case), so the GPU test rebuilds
identical inputs from the same seeds.

Run from the repository root after `sh oracle/build.sh`:
    python tests/golden/make_oracle_curves.py [config1|config2]
"""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from voce_amd import inputs  # noqa: E402
import pyoracle as po  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
MACRO = ("sig_avg", "sig_avg1", "disgrad_avg", "velgrad_avg", "eps_avg", "edot_avg")
SCAL = ("e_vm", "evmp", "d_vm", "dvmp", "s_vm", "s_vm1")
NSAMPLE = 1024


def sample_index(V):
    return np.arange(NSAMPLE, dtype=np.int64) * (V // NSAMPLE)


def make(name, case, n):
    d = tempfile.mkdtemp(prefix="voce_golden_")
    inputs.write_case(d, case)
    grain, _ = inputs.voronoi_grains(case.dims, case.n_grains, case.seed)
    j = po.Job(os.path.join(d, "input"), d, threads=os.cpu_count())
    reps = []
    for k in range(n):
        reps.append(j.increment())
        print(name, "increment", k + 1, "iterations", reps[-1]["iter"], flush=True)
    out = {"iter": np.array([r["iter"] for r in reps]), "iter_total": np.array([r["iter_total"] for r in reps])}
    for key in MACRO:
        out[key] = np.stack([r[key] for r in reps])
    for key in SCAL:
        out[key] = np.array([r[key] for r in reps])
    imax = max(r["iter"] for r in reps)
    for key in ("err_e", "err_s", "svm_it"):
        a = np.full((n, imax), np.nan)
        for k, r in enumerate(reps):
            a[k, :r["iter"]] = r[key]
        out[key] = a
    idx = sample_index(j.nvox)
    out["sample_index"] = idx
    for f in ("sig", "crss", "tranmat"):
        x = j.field(f)
        out["sample_" + f] = x[:, idx]
        if f in ("sig", "crss"):
            out["grain_mean_" + f] = np.stack([x[:, grain == g].mean(axis=1) for g in range(1, case.n_grains + 1)])
    j.close()
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)


def main():
    which = sys.argv[1:] or ["config1", "config2"]
    if "config1" in which:
        make("config1_fp64_50incr", inputs.config_case(1), 50)
    if "config2" in which:
        make("config2_fp64_4incr", inputs.config_case(2), 4)


if __name__ == "__main__":
    main()
