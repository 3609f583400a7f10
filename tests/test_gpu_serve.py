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
