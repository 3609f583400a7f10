"""Batched RVEs on the GPU (include/voce_batch.h): a batch member must give
the same results as the same RVE run alone, the batch must match the oracle
where the reference defines the answer (symmetric L), and a per-increment
velocity gradient must be honoured."""
import os

import numpy as np
import pytest

import voce_amd
from voce_amd import inputs
from gpu_helpers import oracle_run, rel, say, write

pytestmark = pytest.mark.gpu

FIELDS = ("sig", "eps", "crss", "tranmat", "gamtot")


def _specs(skew=True):
    # uneven grid sizes (radix 2 and 3), own seeds and velocity gradients
    dims = [(8, 8, 8), (12, 6, 8), (16, 16, 16), (8, 8, 8), (6, 12, 4)]
    out = []
    for i, d in enumerate(dims):
        rng = inputs.SplitMix64(77 + i)
        L = inputs._traceless_sym(rng)
        L *= 1e-3 / inputs._vm(L)
        if skew:
            W = inputs._skew(rng)
            L = L + W * (1e-3 / inputs._vm(W))
        # 4 significant digits: the same numbers fit a reference namelist line
        L = np.vectorize(lambda x: float("%.4g" % x))(L)
        out.append(inputs.RveSpec(d, 3 + i, 900 + i, L))
    return out


def _run_alone(cases, n, precision, tune_choices=None):
    reps, fields = [], []
    for i, cs in enumerate(cases):
        s = voce_amd.Solver.from_case(cs, precision)
        if tune_choices is not None:
            s.apply_tuning(tune_choices[i])
        r = [s.increment() for _ in range(n)]
        reps.append(r)
        fields.append({f: s.field(f) for f in FIELDS})
        s.close()
    return reps, fields


def _run_batch(cases, n, precision, tune):
    b = voce_amd.Batch(cases, precision, tune=tune)
    choices = [b.member(i).tuning() for i in range(len(cases))]
    reps = [b.increment() for _ in range(n)]
    fields = [{f: b.member(i).field(f) for f in FIELDS} for i in range(len(cases))]
    info = b.info()
    info["choices"] = choices
    b.close()
    return reps, fields, info


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_member_equals_alone(tmp_path, precision):
    say("batch vs alone", precision)
    cases = inputs.load_specs(str(tmp_path), _specs())
    n = 3
    ra, fa = _run_alone(cases, n, precision)
    rb, fb, info = _run_batch(cases, n, precision, tune=False)
    say("batch info", info)
    assert info["members"] == len(cases) and info["groups"] == 4
    for i in range(len(cases)):
        for k in range(n):
            a, b = ra[i][k], rb[k][i]
            assert a["iter"] == b["iter"] and a["istep"] == b["istep"]
            for key in ("sig_avg", "disgrad_avg", "velgrad_avg", "eps_avg", "edot_avg"):
                assert np.array_equal(a[key], b[key]), (i, k, key)
            assert np.array_equal(a["err_s"], b["err_s"]) and np.array_equal(a["err_e"], b["err_e"])
        for f in FIELDS:
            assert np.array_equal(fa[i][f], fb[i][f]), (i, f)


def test_batch_tuned_equals_tuned_alone(tmp_path):
    say("tuned batch vs tuned alone")
    cases = inputs.load_specs(str(tmp_path), _specs())
    n = 2
    rb, fb, info = _run_batch(cases, n, "fp64", tune=True)
    say("group-tuned choices", info["choices"])
    ra, fa = _run_alone(cases, n, "fp64", tune_choices=info["choices"])   # the batch's choices, applied alone
    for i in range(len(cases)):
        for k in range(n):
            assert ra[i][k]["iter"] == rb[k][i]["iter"]
            assert np.array_equal(ra[i][k]["sig_avg"], rb[k][i]["sig_avg"])
        for f in FIELDS:
            assert np.array_equal(fa[i][f], fb[i][f]), (i, f)


def test_batch_matches_oracle_symmetric_L(tmp_path):
    """With W = 0 the per-RVE gradient is expressible in the reference's
    namelist (VelGrad_BC, all flags 1): every member against the oracle."""
    say("batch vs oracle")
    specs = _specs(skew=False)[:3]
    cases = inputs.load_specs(str(tmp_path / "b"), specs)
    n = 3
    rb, fb, _ = _run_batch(cases, n, "fp64", tune=True)
    for i, sp in enumerate(specs):
        L = sp.velgrad
        c = inputs.full_velocity_case(dims=sp.dims)
        c.n_grains, c.seed = sp.n_grains, sp.seed
        c.velgrad = (L[0, 0], L[1, 1], L[2, 2], L[1, 2], L[0, 2], L[0, 1])
        path, base = write(tmp_path / ("o%d" % i), c)
        ro, fo = oracle_run(path, base, n, fields=FIELDS)
        for k in range(n):
            assert ro[k]["iter"] == rb[k][i]["iter"]
            assert rel(rb[k][i]["sig_avg"], ro[k]["sig_avg"]) < 1e-8
        assert rel(fb[i]["sig"], fo["sig"]) < 1e-7
        assert rel(fb[i]["crss"], fo["crss"]) < 1e-9


def test_mixed_batch_matches_oracle_symmetric_L(tmp_path):
    """The MIXED batch (FP32 FFTs and Green operator per member) against the
    FP64 oracle, W = 0: macro stress within 1e-4 at every increment."""
    say("mixed batch vs oracle")
    specs = _specs(skew=False)[:3]
    cases = inputs.load_specs(str(tmp_path / "b"), specs)
    n = 3
    rb, fb, _ = _run_batch(cases, n, "mixed", tune=True)
    for i, sp in enumerate(specs):
        L = sp.velgrad
        c = inputs.full_velocity_case(dims=sp.dims)
        c.n_grains, c.seed = sp.n_grains, sp.seed
        c.velgrad = (L[0, 0], L[1, 1], L[2, 2], L[1, 2], L[0, 2], L[0, 1])
        path, base = write(tmp_path / ("o%d" % i), c)
        ro, fo = oracle_run(path, base, n, fields=("sig",))
        for k in range(n):
            r = rel(rb[k][i]["sig_avg"], ro[k]["sig_avg"])
            say(sp.dims, "incr", k + 1, "mixed batch vs oracle", r)
            assert r < 1e-4
        assert rel(fb[i]["sig"], fo["sig"]) < 1e-5


def test_set_velgrad(tmp_path):
    say("set_velgrad")
    sp = _specs()[2]
    cases = inputs.load_specs(str(tmp_path), [sp, sp])
    a = voce_amd.Solver.from_case(cases[0])
    b = voce_amd.Solver.from_case(cases[1])
    for _ in range(2):
        ra, rb = a.increment(), b.increment()
    # re-imposing the same gradient changes nothing
    b.set_velgrad(sp.velgrad)
    ra, rb = a.increment(), b.increment()
    assert np.array_equal(ra["sig_avg"], rb["sig_avg"]) and ra["iter"] == rb["iter"]
    # a new gradient: under full velocity control the macroscopic velocity
    # gradient of the increment is exactly the imposed one
    rng = inputs.SplitMix64(5)
    L2 = inputs.request_velgrad(rng)
    b.set_velgrad(L2)
    rb = b.increment()
    assert np.abs(rb["velgrad_avg"] - L2).max() < 1e-15
    assert not np.array_equal(ra["sig_avg"], rb["sig_avg"])
    a.close()
    b.close()


def test_batch_radix5_radix7_odd_members(tmp_path):
    """Members of sizes 9^3, 10^3, 15x9x7 and 20^3 in one batch: each equals
    its stand-alone run bit for bit and matches the oracle's iteration count."""
    specs = [inputs.RveSpec(d, 4, 1200 + i, np.zeros((3, 3))) for i, d in
             enumerate([(9, 9, 9), (10, 10, 10), (15, 9, 7), (20, 20, 20)])]
    for sp in specs:
        L = np.diag([1e-3, -0.5e-3, -0.5e-3])
        sp.velgrad = L
    cases = inputs.load_specs(str(tmp_path / "b"), specs)
    n = 2
    ra, fa = _run_alone(cases, n, "fp64")
    rb, fb, info = _run_batch(cases, n, "fp64", tune=False)
    for i, sp in enumerate(specs):
        for k in range(n):
            assert ra[i][k]["iter"] == rb[k][i]["iter"]
            assert np.array_equal(ra[i][k]["sig_avg"], rb[k][i]["sig_avg"])
        for f in FIELDS:
            assert np.array_equal(fa[i][f], fb[i][f]), (i, f)
        c = inputs.full_velocity_case(dims=sp.dims)
        c.n_grains, c.seed = sp.n_grains, sp.seed
        c.velgrad = (1e-3, -0.5e-3, -0.5e-3, 0.0, 0.0, 0.0)
        path, base = write(tmp_path / ("o%d" % i), c)
        ro, fo = oracle_run(path, base, n, fields=("sig",))
        for k in range(n):
            say(sp.dims, "incr", k + 1, "iters batch/oracle", rb[k][i]["iter"], ro[k]["iter"])
            assert rb[k][i]["iter"] == ro[k]["iter"]
            assert rel(rb[k][i]["sig_avg"], ro[k]["sig_avg"]) < 1e-8
        assert rel(fb[i]["sig"], fo["sig"]) < 1e-7


def test_batch_member_with_split_plane_passes(tmp_path):
    """A 3 x 256 x 64 member (planes past one block's LDS: the split z and y
    passes) and a 2 x 256 x 256 one (the line-parallel split passes) batched
    with a 16^3 one: each equals its stand-alone run bit for bit and matches
    the oracle."""
    specs = [inputs.RveSpec(d, 4, 1300 + i, np.zeros((3, 3)))
             for i, d in enumerate([(16, 16, 16), (3, 256, 64), (2, 256, 256)])]
    for sp in specs:
        sp.velgrad = np.diag([1e-3, -0.5e-3, -0.5e-3])
    cases = inputs.load_specs(str(tmp_path / "b"), specs)
    n = 2
    ra, fa = _run_alone(cases, n, "fp64")
    rb, fb, info = _run_batch(cases, n, "fp64", tune=False)
    b = voce_amd.Batch(cases, "fp64", tune=False)
    assert b.member(1).tuning()["plv"] == 5 and b.member(2).tuning()["plv"] == 6
    b.close()
    for i, sp in enumerate(specs):
        for k in range(n):
            assert ra[i][k]["iter"] == rb[k][i]["iter"]
            assert np.array_equal(ra[i][k]["sig_avg"], rb[k][i]["sig_avg"])
        for f in FIELDS:
            assert np.array_equal(fa[i][f], fb[i][f]), (i, f)
        c = inputs.full_velocity_case(dims=sp.dims)
        c.n_grains, c.seed = sp.n_grains, sp.seed
        c.velgrad = (1e-3, -0.5e-3, -0.5e-3, 0.0, 0.0, 0.0)
        path, base = write(tmp_path / ("o%d" % i), c)
        ro, fo = oracle_run(path, base, n, fields=("sig",))
        for k in range(n):
            say(sp.dims, "incr", k + 1, "iters batch/oracle", rb[k][i]["iter"], ro[k]["iter"])
            assert rb[k][i]["iter"] == ro[k]["iter"]
            assert rel(rb[k][i]["sig_avg"], ro[k]["sig_avg"]) < 1e-8
        assert rel(fb[i]["sig"], fo["sig"]) < 1e-7


def test_batch_members_48_point_lines(tmp_path):
    """Members with 48 x 48 planes and Nx = 48 (the 48-point line-parallel
    plane kernels and x pass staged through LDS, tuned per group) batched
    with a 16^3 one: each equals its stand-alone run under the same launch
    choices bit for bit, and matches the oracle."""
    dims = [(4, 48, 48), (48, 8, 12), (16, 16, 16)]
    specs = [inputs.RveSpec(d, 4, 1400 + i, np.diag([1e-3, -0.5e-3, -0.5e-3])) for i, d in enumerate(dims)]
    cases = inputs.load_specs(str(tmp_path / "b"), specs)
    n = 2
    rb, fb, info = _run_batch(cases, n, "fp64", tune=True)
    say("group-tuned choices", info["choices"])
    assert info["choices"][0]["plv"] == 4
    ra, fa = _run_alone(cases, n, "fp64", tune_choices=info["choices"])
    for i, sp in enumerate(specs):
        for k in range(n):
            assert ra[i][k]["iter"] == rb[k][i]["iter"]
            assert np.array_equal(ra[i][k]["sig_avg"], rb[k][i]["sig_avg"])
        for f in FIELDS:
            assert np.array_equal(fa[i][f], fb[i][f]), (i, f)
        c = inputs.full_velocity_case(dims=sp.dims)
        c.n_grains, c.seed = sp.n_grains, sp.seed
        c.velgrad = (1e-3, -0.5e-3, -0.5e-3, 0.0, 0.0, 0.0)
        path, base = write(tmp_path / ("o%d" % i), c)
        ro, fo = oracle_run(path, base, n, fields=("sig",))
        for k in range(n):
            assert rb[k][i]["iter"] == ro[k]["iter"]
            assert rel(rb[k][i]["sig_avg"], ro[k]["sig_avg"]) < 1e-8
        assert rel(fb[i]["sig"], fo["sig"]) < 1e-7
