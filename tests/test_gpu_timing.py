"""Kernel timing (include/voce_prof.h) on the GPU: events bound to the
dispatches of every k-th FFT iteration.  Timing must not change a result,
every sampled iteration contributes one sample per kernel, and the kernels'
own durations fit inside the wall time of the increments they belong to."""
import time

import numpy as np
import pytest

import voce_amd
from voce_amd import inputs
from gpu_helpers import say, write

pytestmark = pytest.mark.gpu


def _run(path, base, n, every, tuning=None):
    s = voce_amd.Solver.from_case(voce_amd.load_case(path, base), "fp64")
    if tuning:
        t = s.tuning()
        t.update(tuning)
        s.apply_tuning(t)
    if every:
        s.prof_enable(True, every=every)
    t0 = time.perf_counter()
    reps = [s.increment() for _ in range(n)]
    wall = time.perf_counter() - t0
    prof = s.prof_read()
    sig = s.field("sig")
    s.close()
    return reps, prof, sig, wall


@pytest.mark.parametrize("tuning", [None, {"plv": 5}])
def test_timing_leaves_results_alone_and_counts_samples(tmp_path, tuning):
    c = inputs.Case(dims=(16, 16, 16), n_grains=8, seed=43, n_steps=4)
    c.iter_max = 20
    path, base = write(tmp_path, c)
    n, every = 4, 3
    r0, p0, s0, _ = _run(path, base, n, 0, tuning)
    r1, p1, s1, wall = _run(path, base, n, every, tuning)
    for a, b in zip(r0, r1):
        assert a["iter"] == b["iter"]
        assert np.array_equal(a["sig_avg"], b["sig_avg"])
    assert np.array_equal(s0, s1)
    assert all(v == (0.0, 0) for v in p0.values())
    # the sampled residue moves by one per increment: iteration it of
    # increment k (k from 0) is timed when (it + k) % every == 0
    want = sum(sum(1 for it in range(r["iter"]) if (it + k) % every == 0) for k, r in enumerate(r1))
    say("samples wanted", want, "timing", p1, "wall s", wall)
    total = 0.0
    for k in ("plane_fwd", "xgreen", "plane_inv", "nr"):
        ms, cnt = p1[k]
        assert cnt == want, k
        assert ms > 0.0, k
        total += ms / cnt
    # every iteration runs the four kernels once
    its = sum(r["iter"] for r in r1)
    assert total * its * 1e-3 < wall
