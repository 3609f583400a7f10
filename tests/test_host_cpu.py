"""CPU checks of the native library: it loads, exports every entry point
include/voce.h declares, and its host-side readers of the reference's file
formats agree with the reference's semantics (and with the oracle's
independent restatement of them).  No compute calls: there is no GPU here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import voce_amd
from voce_amd import inputs
import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "tests", "golden", "ref_inputs")


def header_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for h in sorted(os.listdir(inc)):
        src = open(os.path.join(inc, h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(voce_[a-z_0-9]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = voce_amd.load_library()
    names = header_functions()
    assert set(names) == set(voce_amd.EXPORTS)
    for n in names:
        assert hasattr(lib, n), n


def test_field_ncomp_and_strerror():
    lib = voce_amd.load_library()
    assert [lib.voce_field_ncomp(i) for i in range(18)] == [6, 6, 6, 6, 9, 12, 12, 12, 12, 12, 9, 1, 1, 1, 1, 1, 1, 9]
    assert lib.voce_field_ncomp(99) == -1
    for code in (0, 7, 35, 114, 116, 1139, -1, -2, -3, -4, -5, -6):
        assert len(lib.voce_strerror(code)) >= 2 and lib.voce_strerror(code) != b"unknown error"


def ref_case(tmp_path):
    """The reference's own input, pointed at its shipped 16^3 mesh."""
    txt = open(os.path.join(REF, "input")).read()
    txt = txt.replace("CellDim\t\t64 64 64", "CellDim\t\t16 16 16").replace("mesh_64_rand_125.txt", "mesh_16.txt")
    p = tmp_path / "input16"
    p.write_text(txt)
    return str(p)


def test_load_reference_fixtures(tmp_path):
    path = ref_case(tmp_path)
    cs = voce_amd.load_case(path, REF)
    s = cs.setup
    assert list(s.dims) == [16, 16, 16] and s.n_phases == 2
    assert s.time_step == 0.1 and s.iter_max == 200 and s.err == 5e-5 and cs.n_steps == 4000
    assert list(s.velgrad_flag) == [0, 0, 1, 1, 1, 1] and list(s.stress_flag) == [1, 1, 0, 0, 0, 0]
    assert list(s.velgrad) == [-0.5e-3, -0.5e-3, 1.0e-3, 0, 0, 0]
    p = s.phase[0]
    assert p.type == 0 and list(p.elast) == [169665., 122500., 75827.]
    assert p.nsys == 12 and p.nsrs == 20. and p.gam0 == 1e-3
    assert (p.tau0f, p.tau0b, p.tau1, p.thet0, p.thet1) == (14., 14., 98., 260., 10.)
    assert list(p.bn[0]) == [1, 1, -1] and list(p.bb[0]) == [0, 1, 1]
    assert list(p.bn[11]) == [1, 1, 1] and list(p.bb[11]) == [1, -1, 0]
    assert s.phase[1].type == 1
    H = np.array([list(r) for r in p.hard])
    # "Hirth" and the unterminated "lomer" line are dropped: those entries are 0
    Href = np.zeros(144)
    po.lib().orc_hard_matrix(1.0, 1.0, 1.0, 1.0, 0.0, 0.0, po.ptr(Href))
    assert np.array_equal(H, Href.reshape(12, 12))
    assert H[0, 3] == 0.0 and H[0, 5] == 0.0 and H[0, 1] == 1.0
    e, g, ph = cs.arrays()
    assert e.shape == (4096, 3) and sorted(set(g)) == [1, 2, 3, 4, 5, 6] and set(ph) == {1}
    assert np.allclose(e[0], [113.669, 107.054, 153.016])
    st_or = (C.c_int * 26)()
    po.lib().orc_parse_namelist(path.encode(), st_or, None, 0)
    assert cs.namelist_status == list(st_or)
    cs.free()


@pytest.mark.parametrize("text", [
    "CellDim 4 4\nN_phases 1 2\nN_phases 3\nTimeStep 0.5\n   \nIterMax 7\n",
    "# c\n\nCellDim,4,4,4\nErr 1e-5\nhirth 1 0\nlomer 1 0",
    "CellDim 4 4 4\nUnknown 3\nTimeStep\nN_steps 2 3\n",
])
def test_namelist_semantics_match_oracle(tmp_path, text):
    p = tmp_path / "nl"
    p.write_text(text)
    st_or = (C.c_int * 26)()
    po.lib().orc_parse_namelist(str(p).encode(), st_or, None, 0)
    # the product reader stops before the material file when phases are unset;
    # compare the status vector it leaves
    cs = voce_amd.Case()
    lib = voce_amd.load_library()
    lib.voce_load_case(str(p).encode(), None, C.byref(cs))
    assert list(cs.namelist_status) == list(st_or)


def test_setup_errors_are_codes_not_aborts(tmp_path):
    # a mesh whose record order disagrees with its indices: reference abort 1139
    c = inputs.Case(dims=(4, 4, 4), n_grains=2, seed=3, n_steps=1)
    inputs.write_case(str(tmp_path), c)
    lines = (tmp_path / "mesh.txt").read_text().splitlines()
    lines[5], lines[6] = lines[6], lines[5]
    (tmp_path / "mesh.txt").write_text("\n".join(lines) + "\n")
    with pytest.raises(voce_amd.VoceError) as ei:
        voce_amd.load_case(str(tmp_path / "input"), str(tmp_path))
    assert ei.value.code == 1139
    with pytest.raises(voce_amd.VoceError) as ei:
        voce_amd.load_case(str(tmp_path / "missing"), str(tmp_path))
    assert ei.value.code == -2


def test_generated_case_round_trip(tmp_path):
    c = inputs.Case(dims=(6, 8, 4), n_grains=5, seed=7, n_steps=3)
    g, e = inputs.voronoi_grains(c.dims, c.n_grains, c.seed)
    inputs.write_case(str(tmp_path), c, grain=g, euler=e)
    cs = voce_amd.load_case(str(tmp_path / "input"), str(tmp_path))
    ee, gg, pp = cs.arrays()
    assert np.array_equal(gg, g) and np.allclose(ee, e[g - 1]) and set(pp) == {1}
    assert list(cs.setup.dims) == [6, 8, 4] and cs.n_steps == 3
    assert cs.setup.phase[0].nsrs == 20.0 and cs.setup.phase[0].tau1 == 2.0
    cs.free()


def test_voronoi_generator_properties():
    g, e = inputs.voronoi_grains((8, 8, 8), 6, 42)
    assert g.min() >= 1 and g.max() <= 6
    g2, e2 = inputs.voronoi_grains((8, 8, 8), 6, 42)
    assert np.array_equal(g, g2) and np.array_equal(e, e2)
    # Euler angles in range, Phi from acos -> [0, 180]
    assert np.all((e[:, 0] >= 0) & (e[:, 0] < 360) & (e[:, 1] >= 0) & (e[:, 1] <= 180))
    # periodic: brute-force minimum-image check on a few voxels
    from voce_amd.inputs import SplitMix64
    rng = SplitMix64(42)
    seeds = np.array([[rng.uniform() * 8 for _ in range(3)] for _ in range(6)])
    for v in (0, 77, 300, 511):
        p = np.array([v // 64, (v // 8) % 8, v % 8]) + 0.5
        d = p - seeds
        d -= 8 * np.round(d / 8)
        assert np.argmin((d * d).sum(1)) + 1 == g[v]


def _voronoi_brute(dims, n, seed):
    """Plain-loop restatement of the generator's definition (seeds, then
    Euler triples, then nearest seed by minimum image, ties to the lower id)."""
    import math
    from voce_amd.inputs import SplitMix64
    nx, ny, nz = dims
    rng = SplitMix64(seed)
    seeds = [(rng.uniform() * nx, rng.uniform() * ny, rng.uniform() * nz) for _ in range(n)]
    eul = []
    for _ in range(n):
        a = 360.0 * rng.uniform()
        b = math.acos(2.0 * rng.uniform() - 1.0) * 180.0 / math.pi
        c = 360.0 * rng.uniform()
        eul.append([float("%.6f" % a), float("%.6f" % b), float("%.6f" % c)])
    L = (nx, ny, nz)
    gid = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                p = (i + 0.5, j + 0.5, k + 0.5)
                best, arg = math.inf, 0
                for q in range(n):
                    d = [p[t] - seeds[q][t] for t in range(3)]
                    d = [d[t] - L[t] * round(d[t] / L[t]) for t in range(3)]   # round: ties to even
                    r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
                    if r2 < best:
                        best, arg = r2, q
                gid.append(arg + 1)
    return np.array(gid, dtype=np.int32), np.array(eul)


@pytest.mark.parametrize("dims,n,seed", [((6, 8, 4), 5, 7), ((12, 6, 10), 17, 123456789), ((16, 16, 16), 8, 1)])
def test_voronoi_generator_matches_definition(dims, n, seed):
    g, e = inputs.voronoi_grains(dims, n, seed)
    gb, eb = _voronoi_brute(dims, n, seed)
    assert np.array_equal(g, gb)
    assert np.array_equal(e, eb)


def test_synthetic_case_loader(tmp_path):
    """voce_load_case_voronoi == voce_load_case on the mesh file the same generator writes."""
    import ctypes as C
    c = inputs.Case(dims=(8, 6, 4), n_grains=4, seed=11, n_steps=2)
    g, e = inputs.voronoi_grains(c.dims, c.n_grains, c.seed)
    inputs.write_case(str(tmp_path), c, grain=g, euler=e)
    ref = voce_amd.load_case(str(tmp_path / "input"), str(tmp_path))
    syn = voce_amd.load_case_voronoi(str(tmp_path / "input"), str(tmp_path), (8, 6, 4), 4, 11)
    e1, g1, p1 = ref.arrays()
    e2, g2, p2 = syn.arrays()
    assert np.array_equal(g1, g2) and np.array_equal(e1, e2) and np.array_equal(p1, p2)
    assert list(syn.setup.dims) == [8, 6, 4]
