"""GPU parity: the HIP path (through the C ABI) against the oracle on the
same seeded inputs.  FP64 bar (from the task): macroscopic stress within
1e-8 relative, identical FFT-iteration counts; per-voxel fields within the
tolerances stated at each check.  MIXED bar: macro stress within 1e-4 of FP64."""
import numpy as np
import pytest

import voce_amd
from voce_amd import inputs
from gpu_helpers import gpu_run, oracle_run, rel, say, write

pytestmark = pytest.mark.gpu

FIELDS = ("sig", "eps", "edot", "crss", "tranmat", "gamtot", "gamacum", "disgrad", "svm", "evm", "taux",
          "gamdot", "work_p", "rho_rss", "rho_crss", "work_e", "new_position")

# per-voxel tolerances, relative to the field's largest magnitude
FIELD_TOL = {"sig": 1e-8, "eps": 1e-8, "edot": 1e-7, "crss": 1e-10, "tranmat": 1e-10, "gamtot": 1e-7,
             "gamacum": 1e-8, "disgrad": 1e-9, "svm": 1e-8, "evm": 1e-8, "taux": 1e-7, "gamdot": 1e-7,
             "work_p": 1e-7, "rho_rss": 1e-8, "rho_crss": 1e-10, "work_e": 1e-8, "new_position": 1e-8}


def compare(tmp_path, case, n, grain=None, euler=None, fields=FIELDS, macro_tol=1e-8, tuning=None):
    path, base = write(tmp_path, case, grain, euler)
    say("case", case.dims, case.n_grains, "seed", case.seed, "increments", n, "tuning", tuning)
    g, gf = gpu_run(path, base, n, fields=fields, tuning=tuning)
    o, of = oracle_run(path, base, n, fields=fields)
    for k in range(n):
        say("incr", k + 1, "iters gpu/oracle", g[k]["iter"], o[k]["iter"],
            "sig_avg rel", rel(g[k]["sig_avg"], o[k]["sig_avg"]))
        assert g[k]["iter"] == o[k]["iter"]
        assert g[k]["iter_total"] == o[k]["iter_total"]
        for key in ("sig_avg", "sig_avg1", "disgrad_avg", "velgrad_avg", "eps_avg", "edot_avg"):
            assert rel(g[k][key], o[k][key]) < macro_tol, key
        for key in ("e_vm", "evmp", "d_vm", "dvmp", "s_vm", "s_vm1"):
            assert abs(g[k][key] - o[k][key]) <= macro_tol * max(abs(o[k][key]), 1e-300), key
        # err.out rows: errors are differences of nearby stresses, so looser;
        # an increment that converges in one iteration leaves rows of pure
        # round-off (~1e-16), compared against an absolute floor of 1e-12
        # (far below any tolerance Err a namelist states)
        for key in ("err_s", "err_e"):
            d = np.abs(g[k][key] - o[k][key]).max() if len(o[k][key]) else 0.0
            assert d <= 1e-6 * np.abs(o[k][key]).max(initial=0.0) + 1e-12, key
    for f in fields:
        r = rel(gf[f], of[f])
        say("field", f, "rel", r)
        assert r < FIELD_TOL[f], f
    return g, o


def test_fp64_parity_8cube(tmp_path):
    c = inputs.Case(dims=(8, 8, 8), n_grains=4, seed=5, n_steps=4)
    compare(tmp_path, c, 4)


def test_fp64_parity_16cube_tension(tmp_path):
    # the 16^3 / 8-grain x-tension case with the Cu-like constants
    c = inputs.config_case(1)
    compare(tmp_path, c, 4)


def test_fp64_parity_noncubic_radix3(tmp_path):
    c = inputs.Case(dims=(6, 12, 8), n_grains=5, seed=9, n_steps=3)
    compare(tmp_path, c, 3)


@pytest.mark.parametrize("dims,grains", [((9, 6, 4), 4), ((3, 4, 6), 3), ((2, 2, 2), 2)])
def test_fp64_parity_odd_and_tiny_grids(tmp_path, dims, grains):
    # odd Nx (radix-3 x pass), the smallest grid the kernels take
    c = inputs.Case(dims=dims, n_grains=grains, seed=41, n_steps=2)
    compare(tmp_path, c, 2)


def test_fp64_parity_plane_strain_all_velocity(tmp_path):
    c = inputs.Case(dims=(12, 12, 12), n_grains=6, seed=13, n_steps=3, velgrad_flag=(1, 1, 1, 1, 1, 1),
                    velgrad=(1.0e-3, 0.0, -1.0e-3, 0.0, 0.0, 0.0), stress_flag=(0, 0, 0, 0, 0, 0))
    compare(tmp_path, c, 3)


def test_fp64_parity_shear_stress_bc_and_two_crss(tmp_path):
    # shear stress control (upper-triangle correction) and tau0f != tau0b
    m = inputs.Material(tau0f=16.0, tau0b=20.0)
    c = inputs.Case(dims=(8, 8, 8), n_grains=4, seed=17, n_steps=3, velgrad_flag=(1, 1, 1, 0, 1, 1),
                    velgrad=(1.0e-3, -0.5e-3, -0.5e-3, 0.0, 0.0, 0.0), stress_flag=(0, 0, 0, 1, 0, 0),
                    material=m)
    compare(tmp_path, c, 3)


def test_fp64_parity_gas_voxels(tmp_path):
    c = inputs.Case(dims=(8, 8, 8), n_grains=4, seed=21, n_steps=3)
    g, e = inputs.voronoi_grains(c.dims, c.n_grains, c.seed)
    path, base = write(tmp_path, c, g, e)
    # turn a slab of voxels into phase 2 (Type_phases 1: gas)
    lines = open(base + "/mesh.txt").read().splitlines()
    out = []
    for i, ln in enumerate(lines):
        f = ln.split()
        if i < 64:
            f[7] = "2"
        out.append(" ".join(f))
    open(base + "/mesh.txt", "w").write("\n".join(out) + "\n")
    from gpu_helpers import gpu_run, oracle_run
    gg, gf = gpu_run(path, base, 3, fields=("sig", "crss", "eps"))
    oo, of = oracle_run(path, base, 3, fields=("sig", "crss", "eps"))
    for k in range(3):
        assert gg[k]["iter"] == oo[k]["iter"]
        assert rel(gg[k]["sig_avg"], oo[k]["sig_avg"]) < 1e-8
        assert rel(gg[k]["sig_avg1"], oo[k]["sig_avg1"]) < 1e-8
    assert rel(gf["sig"], of["sig"]) < 1e-8
    assert np.all(gf["sig"][:, :64] == 0.0)


def test_fp64_parity_two_crystal_phases_generic_exponent(tmp_path):
    # Type_phases 0 0: per-voxel phase constants (the multi-phase Newton
    # kernel); phase II uses n = 10, the general integer power
    m2 = inputs.Material(nsrs=10.0, tau0f=24.0, tau0b=24.0, tau1=5.0, theta0=300.0, theta1=40.0)
    c = inputs.Case(dims=(8, 8, 8), n_grains=6, seed=31, n_steps=3, type_phases=(0, 0),
                    elast2=(204600.0, 137700.0, 126200.0, 0.3), material2=m2, phase2_every=2)
    compare(tmp_path, c, 3)


def test_fp64_parity_odd_slip_count(tmp_path):
    # eleven slip systems: the Newton kernel's flow-rule loop takes systems
    # in pairs, so an odd count runs its one-system tail
    m = inputs.Material(slip=list(inputs.FCC_SLIP[:11]))
    c = inputs.Case(dims=(8, 8, 8), n_grains=5, seed=41, n_steps=3, material=m)
    compare(tmp_path, c, 3)


def test_fp64_parity_creep(tmp_path):
    # every component stress-controlled (CREEP_FLAG, init.c:736-740)
    c = inputs.Case(dims=(8, 8, 8), n_grains=4, seed=35, n_steps=3, time_step=0.1,
                    velgrad_flag=(0, 0, 0, 0, 0, 0), velgrad=(0.0,) * 6, stress_flag=(1, 1, 1, 1, 1, 1),
                    stress=(25.0, 0.0, 0.0, 0.0, 0.0, 0.0))
    compare(tmp_path, c, 3)


def test_fp64_parity_no_update_no_hardening(tmp_path):
    c = inputs.Case(dims=(8, 8, 8), n_grains=4, seed=37, n_steps=3, update_flag=0, hard_flag=0)
    compare(tmp_path, c, 3)


def test_noninteger_exponent_aborts_like_reference(tmp_path):
    # pow(r, n - 1) of a negative r is NaN for non-integer n
    # (constitutive.c:40-41): the Newton matrix turns singular and the
    # reference exits with 35; the library returns 35 from the same increment
    c = inputs.Case(dims=(8, 8, 8), n_grains=4, seed=33, n_steps=1, material=inputs.Material(nsrs=12.5))
    path, base = write(tmp_path, c)
    with pytest.raises(RuntimeError, match="code 35"):
        oracle_run(path, base, 1)
    s = voce_amd.Solver.from_case(voce_amd.load_case(path, base))
    with pytest.raises(voce_amd.VoceError) as ei:
        s.increment()
    s.close()
    assert ei.value.code == 35


def test_schmid_field_is_the_next_increments(tmp_path):
    # the Schmid field is update_schmid of the current TranMat, i.e. the
    # reference's Schm_gr once its next increment has refreshed it
    # (evolution.c:320-322, kinematics.c:426-459)
    c = inputs.Case(dims=(8, 8, 8), n_grains=4, seed=43, n_steps=3)
    path, base = write(tmp_path, c)
    g, gf = gpu_run(path, base, 2, fields=("schmid",))
    o, of = oracle_run(path, base, 3, fields=("schmid",))
    assert rel(gf["schmid"], of["schmid"]) < 1e-10


def test_deterministic_repeat(tmp_path):
    c = inputs.Case(dims=(8, 8, 8), n_grains=3, seed=2, n_steps=3)
    path, base = write(tmp_path, c)
    a, af = gpu_run(path, base, 3, fields=("sig", "crss"))
    b, bf = gpu_run(path, base, 3, fields=("sig", "crss"))
    for k in range(3):
        assert np.array_equal(a[k]["sig_avg"], b[k]["sig_avg"])
        assert np.array_equal(a[k]["err_s"], b[k]["err_s"])
    assert np.array_equal(af["sig"], bf["sig"]) and np.array_equal(af["crss"], bf["crss"])


def test_mixed_precision_within_1e4(tmp_path):
    c = inputs.Case(dims=(16, 16, 16), n_grains=8, seed=1, n_steps=4)
    path, base = write(tmp_path, c)
    f64, ff = gpu_run(path, base, 4, "fp64", fields=("sig",))
    mix, mf = gpu_run(path, base, 4, "mixed", fields=("sig",))
    for k in range(4):
        r = rel(mix[k]["sig_avg"], f64[k]["sig_avg"])
        say("incr", k + 1, "mixed vs fp64 macro rel", r, "iters", mix[k]["iter"], f64[k]["iter"])
        assert r < 1e-4
    # per-voxel: FP32 FFT round-off, stated bound 1e-3 of the field maximum
    assert rel(mf["sig"], ff["sig"]) < 1e-3


@pytest.mark.parametrize("dims,tuning", [((16, 16, 16), {"plv": 4}), ((12, 16, 16), {"plv": 4}),
                                         ((32, 32, 32), {"plv": 4}),
                                         ((16, 16, 16), {"plv": 4, "xv": 3, "tk": 4}),
                                         ((16, 16, 16), {"plv": 4, "xv": 3, "tk": 8}),
                                         ((32, 12, 16), {"xv": 3, "tk": 4}),
                                         ((32, 32, 32), {"plv": 4, "xv": 3, "tk": 8}),
                                         ((16, 16, 16), {"plv": 4, "xv": 4, "tk": 8}),
                                         ((16, 16, 16), {"xv": 4, "tk": 4}),
                                         ((32, 12, 16), {"xv": 4, "tk": 8}),
                                         ((32, 12, 16), {"xv": 4, "tk": 4}),
                                         ((16, 16, 16), {"xv": 4, "tk": 2}),
                                         ((32, 12, 16), {"xv": 4, "tk": 2}),
                                         ((32, 32, 32), {"plv": 4, "xv": 4, "tk": 8}),
                                         ((2, 128, 128), {"plv": 4}), ((3, 128, 128), {"plv": 4}),
                                         ((128, 12, 16), {"xv": 4, "tk": 8}), ((128, 12, 16), {"xv": 4, "tk": 2}),
                                         ((128, 8, 8), {"xv": 4, "tk": 4}),
                                         ((4, 48, 48), {"plv": 4}), ((48, 8, 12), {"xv": 4, "tk": 8}),
                                         ((48, 12, 8), {"xv": 4, "tk": 2}), ((48, 8, 8), {"xv": 4, "tk": 4}),
                                         ((256, 8, 8), {"xv": 4, "tk": 4}), ((256, 6, 8), {"xv": 4, "tk": 2})])
def test_fp64_parity_line_parallel(tmp_path, dims, tuning):
    # plane variant 4 (line-parallel kernels of square power-of-two planes,
    # any Nx; 128 x 128 planes take 16-point lines per lane), x-pass
    # variant 3 (line-parallel x pass, Nx = 16, 32, 64, any plane) and
    # x-pass variant 4 (staged through LDS, a whole, half or quarter column
    # block per block, Nx = 16 to 256, two k points per thread at 128, one
    # or two at 48, up to three at 256 with four lines per wave; 12 x 9 =
    # 108, 12 x 5 = 60 and 6 x 5 = 30 columns leave a partial last block),
    # 48-point lines as 4 lanes of 12 points, against the oracle
    c = inputs.Case(dims=dims, n_grains=6, seed=19, n_steps=3)
    compare(tmp_path, c, 3, tuning=tuning)


@pytest.mark.parametrize("tuning", [{"plv": 4, "xv": 3, "tk": 4}, {"plv": 4, "xv": 4, "tk": 16},
                                    {"plv": 4, "xv": 4, "tk": 8}, {"plv": 4, "xv": 4, "tk": 4}])
def test_mixed_line_parallel_planes_against_oracle(tmp_path, tuning):
    c = inputs.Case(dims=(32, 32, 32), n_grains=8, seed=23, n_steps=3)
    path, base = write(tmp_path, c)
    g, gf = gpu_run(path, base, 3, "mixed", fields=("sig",), tuning=tuning)
    o, of = oracle_run(path, base, 3, fields=("sig",))
    for k in range(3):
        r = rel(g[k]["sig_avg"], o[k]["sig_avg"])
        say("incr", k + 1, "mixed (plane variant 4) vs oracle", r)
        assert r < 1e-4
    assert rel(gf["sig"], of["sig"]) < 1e-6


@pytest.mark.parametrize("dims,tuning", [((128, 8, 16), {"xv": 4, "tk": 16}), ((128, 8, 16), {"xv": 4, "tk": 4}),
                                         ((48, 12, 8), {"xv": 4, "tk": 16}), ((48, 12, 8), {"xv": 4, "tk": 8}),
                                         ((4, 48, 48), {"plv": 4}), ((256, 8, 12), {"xv": 4, "tk": 8}),
                                         ((256, 8, 12), {"xv": 4, "tk": 4})])
def test_mixed_staged_x_pass_nx128_nx48_against_oracle(tmp_path, dims, tuning):
    # FP32 x pass staged through LDS at Nx = 128 (two k points per thread)
    # and Nx = 48 (one or two), FP32 line-parallel 48 x 48 planes
    c = inputs.Case(dims=dims, n_grains=6, seed=37, n_steps=2)
    path, base = write(tmp_path, c)
    g, gf = gpu_run(path, base, 2, "mixed", fields=("sig",), tuning=tuning)
    o, of = oracle_run(path, base, 2, fields=("sig",))
    for k in range(2):
        r = rel(g[k]["sig_avg"], o[k]["sig_avg"])
        say("incr", k + 1, "mixed", dims, tuning, "vs oracle", r)
        assert r < 1e-4
    assert rel(gf["sig"], of["sig"]) < 1e-6


def test_mixed_line_parallel_128_planes_against_oracle(tmp_path):
    # FP32 line-parallel plane kernels of 128 x 128 planes
    c = inputs.Case(dims=(2, 128, 128), n_grains=8, seed=29, n_steps=2)
    path, base = write(tmp_path, c)
    g, gf = gpu_run(path, base, 2, "mixed", fields=("sig",), tuning={"plv": 4})
    o, of = oracle_run(path, base, 2, fields=("sig",))
    for k in range(2):
        r = rel(g[k]["sig_avg"], o[k]["sig_avg"])
        say("incr", k + 1, "mixed (plane variant 4, 128 x 128) vs oracle", r)
        assert r < 1e-4
    assert rel(gf["sig"], of["sig"]) < 1e-6


@pytest.mark.parametrize("dims,n", [((10, 10, 10), 3), ((20, 20, 20), 2), ((15, 9, 7), 2), ((14, 21, 10), 2),
                                    ((5, 7, 25), 2)])
def test_fp64_parity_radix5_radix7_odd_planes(tmp_path, dims, n):
    # sizes with factors 5 and 7 and odd Ny / Nz (an odd plane packs its last
    # row with a row of zeros), against the oracle
    c = inputs.Case(dims=dims, n_grains=5, seed=29, n_steps=n)
    compare(tmp_path, c, n)


@pytest.mark.parametrize("dims,tuning", [((48, 48, 48), {"xv": 4, "tk": 8}), ((24, 24, 24), None),
                                         ((96, 32, 32), None), ((80, 16, 16), None)])
def test_fp64_parity_intermediate_sizes(tmp_path, dims, tuning):
    # whole grids between the workloads' sizes against the oracle: 48^3 with
    # the 48-point line-parallel planes and x pass, 24^3 and 80 x 16 x 16
    # (radix 3 and 5) on the Stockham kernels, 96 x 32 x 32 (a radix-3 x pass
    # beside line-parallel 32-point planes)
    c = inputs.Case(dims=dims, n_grains=12, seed=59, n_steps=2)
    compare(tmp_path, c, 2, tuning=tuning)


@pytest.mark.parametrize("dims,tuning", [((48, 48, 48), {"xv": 4, "tk": 16}), ((24, 24, 24), None)])
def test_mixed_intermediate_sizes_against_oracle(tmp_path, dims, tuning):
    c = inputs.Case(dims=dims, n_grains=12, seed=61, n_steps=2)
    path, base = write(tmp_path, c)
    g, gf = gpu_run(path, base, 2, "mixed", fields=("sig",), tuning=tuning)
    o, of = oracle_run(path, base, 2, fields=("sig",))
    for k in range(2):
        r = rel(g[k]["sig_avg"], o[k]["sig_avg"])
        say("incr", k + 1, "mixed", dims, "vs oracle", r)
        assert r < 1e-4
    assert rel(gf["sig"], of["sig"]) < 1e-6


@pytest.mark.parametrize("dims", [(16, 16, 16), (12, 10, 14), (9, 7, 8), (8, 30, 20)])
def test_fp64_parity_split_plane_passes_forced(tmp_path, dims):
    # plane variant 5 (separate z and y launches, each block walking its
    # plane in chunks) on grids the whole-plane kernels also run, against
    # the oracle; 7 is an odd Ny (its last row paired with zeros)
    c = inputs.Case(dims=dims, n_grains=5, seed=31, n_steps=2)
    compare(tmp_path, c, 2, tuning={"plv": 5})


@pytest.mark.parametrize("dims,plv", [((4, 8, 200), 1), ((6, 256, 8), 1), ((256, 6, 4), 1), ((5, 144, 150), 5),
                                      ((3, 256, 64), 5)])
def test_fp64_parity_dimensions_above_128(tmp_path, dims, plv):
    # a dimension beyond 128, default launch choices: planes that still fit
    # one block's LDS run whole (256-point lines included), larger ones (144 x
    # 76 and 256 x 33 half-spectrum points) take the split passes; Nx = 256
    # runs the LDS-tiled x pass with a 2-column tile; 200 = 8 x 5^2,
    # 144 = 16 x 9, 150 = 2 x 3 x 5^2
    c = inputs.Case(dims=dims, n_grains=5, seed=37, n_steps=2)
    path, base = write(tmp_path, c)
    s = voce_amd.Solver.from_case(voce_amd.load_case(path, base), "fp64")
    assert s.tuning()["plv"] == plv
    s.close()
    compare(tmp_path, c, 2, fields=("sig", "eps", "crss", "tranmat", "disgrad", "new_position"))


@pytest.mark.parametrize("dims", [(2, 256, 256), (3, 256, 256)])
def test_fp64_parity_line_parallel_split_passes(tmp_path, dims):
    # plane variant 6 (split z and y launches whose 256-point lines are held
    # by 16 lanes of 16 points), with the diagnostics' fluctuation sums,
    # against the oracle
    c = inputs.Case(dims=dims, n_grains=5, seed=43, n_steps=2)
    compare(tmp_path, c, 2, fields=("sig", "eps", "crss", "tranmat", "disgrad", "new_position"), tuning={"plv": 6})


def test_mixed_line_parallel_split_passes_against_oracle(tmp_path):
    c = inputs.Case(dims=(2, 256, 256), n_grains=5, seed=47, n_steps=2)
    path, base = write(tmp_path, c)
    g, gf = gpu_run(path, base, 2, "mixed", fields=("sig",), tuning={"plv": 6})
    o, of = oracle_run(path, base, 2, fields=("sig",))
    for k in range(2):
        r = rel(g[k]["sig_avg"], o[k]["sig_avg"])
        say("incr", k + 1, "mixed (line-parallel split passes) vs oracle", r)
        assert r < 1e-4
    assert rel(gf["sig"], of["sig"]) < 1e-5


def test_mixed_split_plane_passes_against_oracle(tmp_path):
    c = inputs.Case(dims=(3, 256, 64), n_grains=5, seed=41, n_steps=2)
    path, base = write(tmp_path, c)
    g, gf = gpu_run(path, base, 2, "mixed", fields=("sig",), tuning={"plv": 5})
    o, of = oracle_run(path, base, 2, fields=("sig",))
    for k in range(2):
        r = rel(g[k]["sig_avg"], o[k]["sig_avg"])
        say("incr", k + 1, "mixed (split plane passes) vs oracle", r)
        assert r < 1e-4
    assert rel(gf["sig"], of["sig"]) < 1e-5


def test_fp64_parity_9cube_elastic_converges(tmp_path):
    # all-odd grid: no Nyquist planes, the elastic iteration converges early
    c = inputs.Case(dims=(9, 9, 9), n_grains=4, seed=3, n_steps=2, material=inputs.Material(tau0f=1e4, tau0b=1e4))
    g, o = compare(tmp_path, c, 2, fields=("sig", "disgrad", "tranmat", "crss", "svm", "evm"))
    assert [r["iter"] for r in g] == [34, 32]


def test_mixed_radix5_against_oracle(tmp_path):
    c = inputs.Case(dims=(20, 20, 20), n_grains=8, seed=31, n_steps=2)
    path, base = write(tmp_path, c)
    g, gf = gpu_run(path, base, 2, "mixed", fields=("sig",))
    o, of = oracle_run(path, base, 2, fields=("sig",))
    for k in range(2):
        r = rel(g[k]["sig_avg"], o[k]["sig_avg"])
        say("incr", k + 1, "mixed 20^3 vs oracle", r)
        assert r < 1e-4
    assert rel(gf["sig"], of["sig"]) < 1e-6
