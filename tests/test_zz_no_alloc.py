"""voce_increment_noalloc (include/voce.h) through the C ABI.

It must give what voce_increment gives on a twin context, bit for bit, on
two calls in a row, the second with a smaller RVE (the static workspace
keeps the first call's contents, none of which may leak into the second),
and refuse an RVE whose workspace exceeds the array.  The context's own
workspace must be in use again afterwards (voce_increment still agrees)."""
import ctypes as C

import numpy as np
import pytest

import voce_amd
from voce_amd import inputs
from gpu_helpers import say, write

pytestmark = pytest.mark.gpu

FIELDS = ("sig", "eps", "edot", "crss", "tranmat", "gamtot", "disgrad")
KEYS = ("sig_avg", "sig_avg1", "disgrad_avg", "velgrad_avg", "eps_avg", "edot_avg", "err_e", "err_s", "svm_it")


def noalloc(s):
    r = voce_amd.Report()
    rc = voce_amd.load_library().voce_increment_noalloc(s._h, C.byref(r))
    if rc:
        raise voce_amd.VoceError(rc, "voce_increment_noalloc")
    return voce_amd.report_dict(r)


def twins(tmp, case, precision):
    path, base = write(tmp, case)
    cs = voce_amd.load_case(path, base)
    return voce_amd.Solver.from_case(cs, precision), voce_amd.Solver.from_case(cs, precision), cs


def same(a, b, what):
    assert a["iter"] == b["iter"] and a["istep"] == b["istep"] and a["iter_total"] == b["iter_total"], what
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), (what, k)


def test_noalloc_equals_increment_two_calls_smaller_second(tmp_path):
    # first call: a 16^3 RVE in FP64; second: an 8x8x6 RVE in MIXED (a
    # smaller workspace, and spectra of another precision and layout)
    big = inputs.Case(dims=(16, 16, 16), n_grains=8, seed=1, n_steps=3)
    small = inputs.Case(dims=(8, 8, 6), n_grains=4, seed=7, n_steps=3)
    ra, rb, ca = twins(tmp_path / "big", big, "fp64")
    sa, sb, cb = twins(tmp_path / "small", small, "mixed")
    say("no-alloc: 16^3 fp64 then 8x8x6 mixed, 3 rounds")
    for k in range(3):
        same(ra.increment(), noalloc(rb), ("16^3", k))
        same(sa.increment(), noalloc(sb), ("8x8x6", k))
    for f in FIELDS:
        assert np.array_equal(ra.field(f), rb.field(f)), ("16^3", f)
        assert np.array_equal(sa.field(f), sb.field(f)), ("8x8x6", f)
    # the contexts' own workspace again afterwards
    same(ra.increment(), rb.increment(), "16^3 after")
    same(sa.increment(), sb.increment(), "8x8x6 after")
    for s in (ra, rb, sa, sb):
        s.close()


def test_noalloc_refuses_oversize(tmp_path):
    # 96^3 in FP64 needs about 250 MB of workspace, more than the array holds
    c = inputs.Case(dims=(8, 8, 8), n_grains=4, seed=3, n_steps=1)
    path, base = write(tmp_path, c)
    cs = voce_amd.load_case_voronoi(path, base, (96, 96, 96), 64, 9)
    s = voce_amd.Solver.from_case(cs, "fp64")
    with pytest.raises(voce_amd.VoceError) as e:
        noalloc(s)
    assert e.value.code == -7
    # the context is untouched and still steps through its own workspace
    r = s.increment()
    assert r["istep"] == 1 and r["iter"] >= 1
    s.close()
