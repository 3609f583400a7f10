"""Helpers shared by the GPU parity tests: run one synthetic case through the
HIP library (via the C ABI) and through the oracle, increment by increment."""
import os
import sys

import numpy as np

import voce_amd
from voce_amd import inputs
import pyoracle as po

THREADS = max(1, min(16, os.cpu_count() or 1))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def write(tmp, case, grain=None, euler=None):
    d = str(tmp)
    inputs.write_case(d, case, grain=grain, euler=euler)
    return os.path.join(d, "input"), d


def gpu_run(path, base, n, precision="fp64", diag=True, fields=(), tuning=None):
    cs = voce_amd.load_case(path, base)
    s = voce_amd.Solver.from_case(cs, precision)
    if tuning:
        t = s.tuning()
        t.update(tuning)
        s.apply_tuning(t)
    if diag:
        s.set_diagnostics(True)
    reps = [s.increment() for _ in range(n)]
    out = {f: s.field(f) for f in fields}
    s.close()
    return reps, out


def oracle_run(path, base, n, fields=()):
    j = po.Job(path, base, threads=THREADS)
    reps = [j.increment() for _ in range(n)]
    out = {f: j.field(f) for f in fields}
    j.close()
    return reps, out


def rel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
