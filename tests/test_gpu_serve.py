"""Request server on the GPU (include/voce_batch.h): each served increment
gives the same result alone and in a mix, matches the same increments on a
stand-alone context, and an FP64 audit reports the mixed-precision error
without disturbing the served result."""
import numpy as np
import pytest

import voce_amd
from voce_amd import inputs
from gpu_helpers import say

pytestmark = pytest.mark.gpu

DIMS = [(8, 8, 8), (16, 16, 16), (12, 6, 8), (6, 12, 4), (8, 8, 8)]
KEYS = ("sig_avg", "disgrad_avg", "velgrad_avg", "eps_avg", "edot_avg")


def _cases(tmp):
    specs = [inputs.RveSpec(d, 3 + i, 7000 + i, np.zeros((3, 3))) for i, d in enumerate(DIMS)]
    return inputs.load_specs(str(tmp), specs)


def _requests(n, seed=11, audit_every=0):
    rng = inputs.SplitMix64(seed)
    out = []
    for t in range(n):
        rid = int(rng.uniform() * len(DIMS))
        L = inputs.request_velgrad(rng)
        out.append((rid, L, bool(audit_every) and t % audit_every == 0, t))
    return out


def _server(cases, **kw):
    s = voce_amd.Server(len(cases), "mixed", **kw)
    for i, cs in enumerate(cases):
        s.add_rve(i, cs)
    return s


def _by_tag(results):
    return {r["tag"]: r for r in results}


def _same(a, b):
    assert a["iter"] == b["iter"] and a["istep"] == b["istep"], (a["tag"], a["iter"], b["iter"])
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), (a["tag"], k)


def test_alone_equals_mix(tmp_path):
    cases = _cases(tmp_path)
    reqs = _requests(24)
    say("alone")
    sa = _server(cases)
    alone = []
    for rid, L, au, tag in reqs:
        sa.submit(rid, L, au, tag)
        alone += sa.drain()
    fa = {i: sa.rve(i).field("sig") for i in range(len(cases))}
    say("mix")
    sb = _server(cases)
    for rid, L, au, tag in reqs:
        sb.submit(rid, L, au, tag)
    mix = sb.drain()
    fb = {i: sb.rve(i).field("sig") for i in range(len(cases))}
    info = sb.info()
    say("info", info)
    assert len(alone) == len(mix) == len(reqs)
    A, B = _by_tag(alone), _by_tag(mix)
    for t in A:
        _same(A[t], B[t])
    for i in fa:
        assert np.array_equal(fa[i], fb[i])
    sa.close()
    sb.close()


def test_served_equals_context(tmp_path):
    """The served increments of one id equal the same increments run on a
    stand-alone context with voce_set_velgrad before each."""
    cases = _cases(tmp_path)
    reqs = [r for r in _requests(40) if r[0] == 1][:5]
    s = _server(cases)
    for rid, L, au, tag in reqs:
        s.submit(rid, L, False, tag)
    served = sorted(s.drain(), key=lambda r: r["tag"])
    c = voce_amd.Solver.from_case(cases[1], "mixed")
    c.apply_tuning(s.rve(1).tuning())   # the server's launch choice: the same kernels
    for (rid, L, _, tag), r in zip(reqs, served):
        c.set_velgrad(L)
        _same(c.increment() | {"tag": tag}, r)
    assert np.array_equal(c.field("crss"), s.rve(1).field("crss"))
    c.close()
    s.close()


def test_audit(tmp_path):
    cases = _cases(tmp_path)
    reqs = _requests(20, seed=3)
    s1 = _server(cases)
    for rid, L, _, tag in reqs:
        s1.submit(rid, L, tag % 4 == 0, tag)
    r1 = _by_tag(s1.drain())
    s2 = _server(cases)
    for rid, L, _, tag in reqs:
        s2.submit(rid, L, False, tag)
    r2 = _by_tag(s2.drain())
    n_aud = 0
    for t in r1:
        _same(r1[t], r2[t])     # the audit leaves the served result alone
        if r1[t]["audited"]:
            n_aud += 1
            say("audit", t, r1[t]["audit_sig_rel"])
            assert 0.0 < r1[t]["audit_sig_rel"] < 1e-4
    assert n_aud == 5
    s1.close()
    s2.close()


def test_admission_under_memory_pressure(tmp_path):
    cases = _cases(tmp_path)
    reqs = _requests(16, seed=5)
    free = _server(cases)
    for rid, L, au, tag in reqs:
        free.submit(rid, L, au, tag)
    rf = _by_tag(free.drain())
    # room for about one 16^3 workspace (about 1 MB) at a time, at most two in flight
    tight = _server(cases, max_inflight=2, workspace_bytes=1_300_000)
    for rid, L, au, tag in reqs:
        tight.submit(rid, L, au, tag)
    rt = _by_tag(tight.drain())
    assert tight.info()["pool_used"] == 0
    for t in rf:
        _same(rf[t], rt[t])
    free.close()
    tight.close()
    # a pool smaller than one 16^3 workspace: refused at submission, not queued for ever
    tiny = _server(cases, workspace_bytes=300_000)
    with pytest.raises(voce_amd.VoceError) as e:
        tiny.submit(1, reqs[0][1], False, 0)
    assert e.value.code == -5
    tiny.submit(0, reqs[0][1], False, 1)   # an 8^3 workspace fits
    assert [r["tag"] for r in tiny.drain()] == [1]
    tiny.close()


def test_audit_leaves_diagnostics_alone(tmp_path):
    """With diagnostics on for a served id, an audited request's FP64 shadow
    must not touch the id's arrays (its fluctuation sums feed new_position)."""
    cases = _cases(tmp_path)
    reqs = [r for r in _requests(40, seed=9) if r[0] == 2][:3]
    out = []
    for audit in (True, False):
        s = _server(cases)
        s.rve(2).set_diagnostics(True)
        for rid, L, _, tag in reqs:
            s.submit(rid, L, audit, tag)
            s.drain()
        out.append({f: s.rve(2).field(f) for f in ("new_position", "sig", "svm", "work_p")})
        s.close()
    for f in out[0]:
        assert np.array_equal(out[0][f], out[1][f]), f
    assert np.abs(out[0]["new_position"]).max() > 0


def test_audits_sharing_a_tag(tmp_path):
    """Tags are the caller's and need not be unique: two audited requests on
    different ids in flight together under tag 0 each get their own result
    and their own audit."""
    cases = _cases(tmp_path)
    rng = inputs.SplitMix64(21)
    La, Lb = inputs.request_velgrad(rng), inputs.request_velgrad(rng)
    s = _server(cases)
    s.submit(0, La, True, 0)
    s.submit(3, Lb, True, 0)
    s.submit(4, La, False, 0)
    res = s.drain()
    s.close()
    assert len(res) == 3 and sorted(r["id"] for r in res) == [0, 3, 4]
    ref = _server(cases)
    for rid, L in ((0, La), (3, Lb), (4, La)):
        ref.submit(rid, L, False, 7)
    want = {r["id"]: r for r in ref.drain()}
    ref.close()
    for r in res:
        w = want[r["id"]]
        assert r["iter"] == w["iter"] and np.array_equal(r["sig_avg"], w["sig_avg"]), r["id"]
        if r["id"] in (0, 3):
            assert r["audited"] and 0.0 < r["audit_sig_rel"] < 1e-4
        else:
            assert not r["audited"]


def test_shadow_allocation_failure_leaves_server_consistent(tmp_path):
    """An audit whose FP64 shadow cannot be allocated fails that scheduling
    round: the request goes back in its place with nothing held, what was
    admitted before it runs on, and a later round serves it normally."""
    cases = _cases(tmp_path)
    rng = inputs.SplitMix64(13)
    # one request per id, the third audited: ids 0 and 1 are admitted before it
    plan = [(i, inputs.request_velgrad(rng), i == 2, 40 + i) for i in range(len(DIMS))]
    s = _server(cases)
    s.inject_fault(1, 1)
    for rid, L, au, tag in plan:
        s.submit(rid, L, au, tag)
    with pytest.raises(voce_amd.VoceError) as e:
        s.step()
    assert e.value.code == -5
    info = s.info()
    say("after the failed round", info)
    assert info["queued"] == 3 and info["inflight"] == 2 and info["audits"] == 0
    got = _by_tag(s.drain() + s.poll())
    info = s.info()
    assert info["queued"] == 0 and info["inflight"] == 0 and info["pool_used"] == 0
    s.close()
    ref = _server(cases)
    for rid, L, au, tag in plan:
        ref.submit(rid, L, au, tag)
    want = _by_tag(ref.drain())
    ref.close()
    assert sorted(got) == sorted(want)
    for t in want:
        _same(got[t], want[t])
        assert got[t]["audited"] == want[t]["audited"]


def _voigt_L(setup):
    v = setup.velgrad
    return np.array([[v[0], v[5], v[4]], [v[5], v[1], v[3]], [v[4], v[3], v[2]]])


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-8), ("mixed", 1e-4)])
def test_served_increments_match_oracle(tmp_path, precision, tol):
    """Served increments of 16^3 and 32^3 ids against the oracle: each id's
    requests carry its namelist's symmetric velocity gradient (W = 0, the
    only form VelGrad_BC can state, kinematics.c:5-12), so every served
    increment is an increment of the reference run of that namelist."""
    from gpu_helpers import oracle_run, rel, write
    n = 3
    specs = [((16, 16, 16), 8, 51, (1.0e-3, -0.4e-3, -0.6e-3, 0.2e-3, 0.0, -0.3e-3)),
             ((32, 32, 32), 12, 52, (-0.5e-3, 1.0e-3, -0.5e-3, 0.0, 0.4e-3, 0.1e-3))]
    srv = voce_amd.Server(len(specs), precision, tune=False)
    runs = []
    for i, (dims, grains, seed, vg) in enumerate(specs):
        c = inputs.full_velocity_case(dims=dims, n_steps=n)
        c.n_grains, c.seed, c.velgrad = grains, seed, vg
        path, base = write(tmp_path / ("id%d" % i), c)
        cs = voce_amd.load_case(path, base)
        srv.add_rve(i, cs)
        runs.append((path, base, _voigt_L(cs.setup), cs))
    res = []
    for k in range(n):
        for i in range(len(specs)):
            srv.submit(i, runs[i][2], False, 10 * k + i)
        res += srv.drain()
    by = _by_tag(res)
    sig = [srv.rve(i).field("sig") for i in range(len(specs))]
    srv.close()
    for i, (path, base, L, cs) in enumerate(runs):
        ro, fo = oracle_run(path, base, n, fields=("sig",))
        for k in range(n):
            r = by[10 * k + i]
            e = rel(r["sig_avg"], ro[k]["sig_avg"])
            say(precision, "id", i, "incr", k + 1, "iters", r["iter"], ro[k]["iter"], "macro rel", e)
            assert e < tol
            if precision == "fp64":
                assert r["iter"] == ro[k]["iter"]
                assert rel(r["disgrad_avg"], ro[k]["disgrad_avg"]) < tol
        fe = rel(sig[i], fo["sig"])
        say(precision, "id", i, "per-voxel stress (max abs / max)", fe)
        assert fe < (1e-7 if precision == "fp64" else 1e-3)


def test_alone_equals_mix_config5_sizes(tmp_path):
    """Ids of the served workload's sizes (16^3, 32^3, 64^3), MIXED: each
    result is the same alone and in the mix."""
    specs = [inputs.RveSpec(d, g, 5_000_000 + i, np.zeros((3, 3)))
             for i, (d, g) in enumerate([((16, 16, 16), 8), ((32, 32, 32), 20), ((64, 64, 64), 40),
                                         ((32, 32, 32), 64), ((16, 16, 16), 30)])]
    cases = inputs.load_specs(str(tmp_path), specs)
    rng = inputs.SplitMix64(55)
    reqs = []
    for t in range(12):
        rid = int(rng.uniform() * len(specs))
        reqs.append((rid, inputs.request_velgrad(rng), t % 5 == 0, t))
    alone = []
    sa = voce_amd.Server(len(cases), "mixed")
    for i, cs in enumerate(cases):
        sa.add_rve(i, cs)
    for rid, L, au, tag in reqs:
        sa.submit(rid, L, au, tag)
        alone += sa.drain()
    sa.close()
    sb = voce_amd.Server(len(cases), "mixed")
    for i, cs in enumerate(cases):
        sb.add_rve(i, cs)
    for rid, L, au, tag in reqs:
        sb.submit(rid, L, au, tag)
    mix = sb.drain()
    sb.close()
    A, B = _by_tag(alone), _by_tag(mix)
    assert len(A) == len(B) == len(reqs)
    for t in A:
        _same(A[t], B[t])
        if A[t]["audited"]:
            assert 0.0 < A[t]["audit_sig_rel"] < 1e-4 and A[t]["audit_sig_rel"] == B[t]["audit_sig_rel"]


def test_fp64_server_with_audits(tmp_path):
    """An FP64 server may audit too: the FP64 shadow runs the default launch
    choices beside tuned rows of the same grid and precision (separate launch
    sets), and agrees with the served result to round-off."""
    cases = _cases(tmp_path)
    reqs = _requests(10, seed=17)
    s = voce_amd.Server(len(cases), "fp64", tune=True)
    for i, cs in enumerate(cases):
        s.add_rve(i, cs)
    for rid, L, _, tag in reqs:
        s.submit(rid, L, tag % 2 == 0, tag)
    res = s.drain()
    s.close()
    assert len(res) == len(reqs)
    aud = [r["audit_sig_rel"] for r in res if r["audited"]]
    say("fp64 audits", aud)
    assert len(aud) == 5 and max(aud) < 1e-10


@pytest.mark.parametrize("dims,plv", [((3, 256, 64), 5), ((2, 256, 256), 6)])
def test_served_split_plane_id_equals_context(tmp_path, dims, plv):
    """An id whose planes take the split z and y passes (3 x 256 x 64; 256 x
    256 planes: the line-parallel ones) served beside a 16^3 id: its
    increments equal the same increments on a stand-alone context, and FP64
    audits of it report the mixed error."""
    specs = [inputs.RveSpec(d, 4, 7100 + i, np.zeros((3, 3))) for i, d in enumerate([dims, (16, 16, 16)])]
    cases = inputs.load_specs(str(tmp_path), specs)
    s = _server(cases)
    assert s.rve(0).tuning()["plv"] == plv
    rng = inputs.SplitMix64(29)
    reqs = []
    for t in range(8):
        reqs.append((t % 2, inputs.request_velgrad(rng), t % 4 == 0, t))
    for rid, L, au, tag in reqs:
        s.submit(rid, L, au, tag)
    served = _by_tag(s.drain())
    c = voce_amd.Solver.from_case(cases[0], "mixed")
    c.apply_tuning(s.rve(0).tuning())
    for rid, L, au, tag in reqs:
        if rid != 0:
            continue
        c.set_velgrad(L)
        r = served[tag]
        _same(c.increment() | {"tag": tag}, r)
        if au:
            say("audit", tag, r["audited"], r["audit_sig_rel"])
            assert r["audited"] and r["audit_sig_rel"] < 1e-4
    assert np.array_equal(c.field("sig"), s.rve(0).field("sig"))
    c.close()
    s.close()
