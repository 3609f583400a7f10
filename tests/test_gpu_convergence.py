"""The convergence exit of the FFT iteration (evolution.c:342) on the GPU.

Most plastic Cu increments run to IterMax (the stress error decays slowly
and, on grids with even sizes, the Nyquist-plane Green operator of
init.c:661-667 keeps a floor under it), so these cases are chosen to stop
early, each at its own count, and are compared with the oracle:

* an isotropic-elastic polycrystal (C11 - C12 = 2 C44): 1, 23, 27 iterations;
* a Cu single crystal (uniform fields): 32, 30, 30, 30;
* elastic Cu polycrystals (CRSS 1e4 MPa) on a 3x4x6 grid (57, 62, 55) and a
  9x8x8 grid (97, 100, 84);
* a batch mixing them with a member that runs to IterMax (converged members
  leave the work while the others iterate), and
* a server round whose rows finish at different iterations (the chunk
  length follows each id's previous count).

FP64 bar: identical iteration counts, macro stress within 1e-8 relative.
"""
import os

import numpy as np
import pytest

import voce_amd
from voce_amd import inputs
from gpu_helpers import oracle_run, rel, say, write
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu

ELASTIC_CU = inputs.Material(tau0f=1e4, tau0b=1e4)
ISO = (200000.0, 100000.0, 50000.0, 0.3)
# fields that stay meaningful when the flow rule is switched off by a huge
# CRSS (the plastic rates are ~(tau/1e4)^20: pure rounding noise)
ELASTIC_FIELDS = ("sig", "disgrad", "tranmat", "crss", "svm", "evm")
UNIFORM_FIELDS = ("sig", "eps", "edot", "crss", "tranmat", "gamtot", "gamacum", "disgrad", "svm", "evm", "taux",
                  "gamdot", "work_p", "rho_rss", "rho_crss", "work_e")


def iso_case(n_steps=3):
    return inputs.Case(dims=(8, 8, 8), n_grains=4, seed=11, n_steps=n_steps, material=ELASTIC_CU, elast=ISO)


def single_crystal(tmp, n_steps=4):
    c = inputs.Case(dims=(8, 8, 8), n_grains=1, n_steps=n_steps)
    g = np.ones(c.nvox, dtype=np.int32)
    e = np.array([[30.0, 40.0, 50.0]])
    return c, g, e


def elastic_cu(dims, n_steps=3):
    return inputs.Case(dims=dims, n_grains=4 if dims[0] > 3 else 3, seed=3, n_steps=n_steps, material=ELASTIC_CU)


def _counts(reps):
    return [r["iter"] for r in reps]


def test_isotropic_elastic_polycrystal_converges(tmp_path):
    g, o = compare(tmp_path, iso_case(), 3, fields=ELASTIC_FIELDS)
    say("iterations", _counts(g))
    assert _counts(g) == [1, 23, 27]


def test_single_crystal_converges(tmp_path):
    c, gr, eu = single_crystal(tmp_path)
    g, o = compare(tmp_path, c, 4, grain=gr, euler=eu, fields=UNIFORM_FIELDS)
    say("iterations", _counts(g))
    assert _counts(g) == [32, 30, 30, 30]


@pytest.mark.parametrize("dims,want", [((3, 4, 6), [57, 62, 55]), ((9, 8, 8), [97, 100, 84])])
def test_elastic_polycrystal_converges(tmp_path, dims, want):
    g, o = compare(tmp_path, elastic_cu(dims), 3, fields=ELASTIC_FIELDS)
    say("iterations", _counts(g))
    assert _counts(g) == want


def _load(tmp, case, grain=None, euler=None):
    path, base = write(tmp, case, grain, euler)
    return path, base, voce_amd.load_case(path, base)


def _mixed_set(tmp_path):
    """Cases that stop at different iteration counts, plus one that runs to
    IterMax (plastic 8^3 x-tension)."""
    sc, gr, eu = single_crystal(tmp_path, 3)
    specs = [("iso", iso_case(), None, None), ("xtal", sc, gr, eu),
             ("el346", elastic_cu((3, 4, 6)), None, None),
             ("plastic", inputs.Case(dims=(8, 8, 8), n_grains=4, seed=5, n_steps=3), None, None)]
    out = []
    for name, c, g, e in specs:
        path, base, cs = _load(tmp_path / name, c, g, e)
        out.append((name, path, base, cs))
    return out


def test_batch_members_converge_at_their_own_counts(tmp_path):
    n = 3
    sets = _mixed_set(tmp_path)
    cases = [cs for _, _, _, cs in sets]
    b = voce_amd.Batch(cases, "fp64", tune=False)
    rb = [b.increment() for _ in range(n)]
    fb = [b.member(i).field("sig") for i in range(len(cases))]
    info = b.info()
    b.close()
    say("batch info", info)
    seen = set()
    for i, (name, path, base, cs) in enumerate(sets):
        ro, fo = oracle_run(path, base, n, fields=("sig",))
        s = voce_amd.Solver.from_case(cs, "fp64")
        ra = [s.increment() for _ in range(n)]
        fa = s.field("sig")
        s.close()
        for k in range(n):
            say(name, "incr", k + 1, "iters batch/alone/oracle", rb[k][i]["iter"], ra[k]["iter"], ro[k]["iter"])
            assert rb[k][i]["iter"] == ra[k]["iter"] == ro[k]["iter"]
            assert np.array_equal(rb[k][i]["sig_avg"], ra[k]["sig_avg"])
            assert np.array_equal(rb[k][i]["err_s"], ra[k]["err_s"])
            assert rel(rb[k][i]["sig_avg"], ro[k]["sig_avg"]) < 1e-8
            seen.add(ro[k]["iter"])
        assert np.array_equal(fb[i], fa)
    # converged members left the work: fewer iterations launched than
    # members x IterMax would be, and several counts below the cap
    assert len([x for x in seen if x < 100]) >= 6


def test_server_rows_finish_at_different_iterations(tmp_path):
    """Each id serves its own namelist's velocity gradient (symmetric: the
    requests reproduce the reference run of that namelist increment by
    increment); rows of one round finish at different iterations."""
    n = 3
    sets = _mixed_set(tmp_path)
    srv = voce_amd.Server(len(sets), "fp64", tune=False)
    Ls = []
    for i, (name, path, base, cs) in enumerate(sets):
        srv.add_rve(i, cs)
        v = cs.setup.velgrad
        Ls.append(np.array([[v[0], v[5], v[4]], [v[5], v[1], v[3]], [v[4], v[3], v[2]]]))
    res = []
    for k in range(n):
        for i in range(len(sets)):
            srv.submit(i, Ls[i], False, 100 * k + i)
        res += srv.drain()
    info = srv.info()
    srv.close()
    say("server info", info)
    by = {r["tag"]: r for r in res}
    for i, (name, path, base, cs) in enumerate(sets):
        ro, _ = oracle_run(path, base, n)
        for k in range(n):
            r = by[100 * k + i]
            say(name, "incr", k + 1, "iters served/oracle", r["iter"], ro[k]["iter"])
            assert r["iter"] == ro[k]["iter"]
            assert rel(r["sig_avg"], ro[k]["sig_avg"]) < 1e-8
            assert rel(r["disgrad_avg"], ro[k]["disgrad_avg"]) < 1e-8
