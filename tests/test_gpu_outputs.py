"""Host driver and output formats on the GPU: `voce_amd/bin/voce input`
(the reference's main, evp.c:7-32) against the oracle's writers on the same
case.  Same file set and names; .iout doubles within the field tolerances of
the parity tests; the int fields of the 104-byte (slip/taux/crss) and 48-byte
(tex) records identical; vm.out / err.out rows equal to their printed
precision (evolution.c:302, 427, 563-564)."""
import glob
import os
import shutil
import subprocess

import numpy as np
import pytest

from voce_amd import inputs
import pyoracle as po
from gpu_helpers import THREADS, say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
VOCE = os.path.join(ROOT, "voce_amd", "bin", "voce")
ORACLE = os.path.join(ROOT, "oracle", "evp_oracle")


def _rows(path):
    out = []
    for ln in open(path).read().splitlines()[1:]:
        if ln.startswith("STEP"):
            out.append(ln)
        else:
            out.append([float(x) for x in ln.split()])
    return out


def _close_rows(a, b, rtol):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        if isinstance(x, str):
            assert x == y
            continue
        x, y = np.array(x), np.array(y)
        assert np.all(np.abs(x - y) <= rtol * np.maximum(np.abs(y), 1e-300) + 1e-300), (x, y)


def test_driver_outputs_match_oracle(tmp_path):
    c = inputs.Case(dims=(8, 8, 8), n_grains=3, seed=21, n_steps=3, print_control=(1, 1), tex_flag=1)
    a, b = tmp_path / "gpu", tmp_path / "oracle"
    inputs.write_case(str(a), c)
    shutil.copytree(str(a), str(b))
    say("running", VOCE)
    r = subprocess.run([VOCE, "input"], cwd=str(a), capture_output=True, text=True, timeout=120)
    say(r.stderr.strip())
    assert r.returncode == 0
    # oracle: the same driver loop for vm.out / err.out, the oracle's writers for the fields
    r = subprocess.run([ORACLE, "input", str(THREADS)], cwd=str(b), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    j = po.Job(os.path.join(str(b), "input"), str(b), threads=THREADS)
    for s in range(1, c.n_steps + 1):
        j.increment()
        j.write_outputs(str(b), s, 1)
    work_e = j.field("work_e").ravel()
    j.close()
    # work_e is written only at the end (evolution.c:603)
    we = np.frombuffer((a / "work_e_final_S000000.iout").read_bytes(), dtype=np.float64)
    assert we.shape == work_e.shape
    assert np.abs(we - work_e).max() <= 1e-7 * max(np.abs(work_e).max(), 1e-300)
    # vm.out: %e (7 significant digits); err.out rows hold differences of nearby stresses
    _close_rows(_rows(a / "vm.out"), _rows(b / "vm.out"), 2e-6)
    _close_rows(_rows(a / "err.out"), _rows(b / "err.out"), 1e-4)
    step_files = sorted(os.path.basename(p) for p in glob.glob(str(b / "*_S*.iout")))
    assert step_files
    V = c.nvox
    for name in step_files:
        fa, fb = a / name, b / name
        assert fa.exists(), name
        ba, bb = fa.read_bytes(), fb.read_bytes()
        assert len(ba) == len(bb), name
        if name.startswith(("slip_sys", "taux_sys", "crss_sys")):
            ra = np.frombuffer(ba, dtype=np.uint8).reshape(V, 104)
            rb = np.frombuffer(bb, dtype=np.uint8).reshape(V, 104)
            assert np.array_equal(ra[:, :8], rb[:, :8]), name          # grain id + padding
            da = ra[:, 8:].copy().view(np.float64)
            db = rb[:, 8:].copy().view(np.float64)
            assert np.abs(da - db).max() <= 1e-7 * max(np.abs(db).max(), 1e-300), name
        elif name.startswith("tex"):
            ra = np.frombuffer(ba, dtype=np.uint8).reshape(V, 48)
            rb = np.frombuffer(bb, dtype=np.uint8).reshape(V, 48)
            assert np.array_equal(ra[:, 24:], rb[:, 24:]), name        # i, j, k, grain, phase, 0
            ea = ra[:, :24].copy().view(np.float64)
            eb = rb[:, :24].copy().view(np.float64)
            assert np.abs(ea - eb).max() < 1e-6, name                 # degrees
        else:
            da = np.frombuffer(ba, dtype=np.float64)
            db = np.frombuffer(bb, dtype=np.float64)
            assert np.abs(da - db).max() <= 1e-7 * max(np.abs(db).max(), 1e-300), name
    # after the last increment the *_final dumps hold the same state as its step dumps
    last = [n for n in step_files if n.endswith(("_S%04d.iout" % c.n_steps, "_S%06d.iout" % c.n_steps))]
    assert last
    for name in last:
        fin = a / _final_name(name)
        assert fin.exists(), fin.name
        assert fin.read_bytes() == (a / name).read_bytes(), name


def _final_name(name):
    """Name of the end-of-run dump (evolution.c:594-611) for a step dump."""
    import re
    prefix, rest = name.rsplit("_S", 1)
    zero = "0" * (len(rest) - len(".iout"))
    m = re.match(r"(eps|sig|els|edot)_(\d)$", prefix)
    if m:
        return "%s_final_%s_S%s.iout" % (m.group(1), m.group(2), zero)
    prefix = {"slip_sys": "slip", "taux_sys": "taux", "crss_sys": "crss"}.get(prefix, prefix)
    return "%s_final_S%s.iout" % (prefix, zero)
