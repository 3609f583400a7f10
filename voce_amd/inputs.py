"""Input files in the reference's own formats.

Writes the three inputs the reference reads -- the namelist (io.c:25-122),
the positional ``.sx`` material file (init.c:51-187) and the 8-column voxel
mesh (init.c:393-423) -- for seeded synthetic polycrystals:

* periodic Voronoi grains: seeds uniform in [0, N)^3, voxel centres at
  (i+0.5, j+0.5, k+0.5), minimum-image distance, ties to the lowest seed id;
* one Bunge Euler triple per grain, uniform on SO(3):
  phi1 = 360 u, Phi = acos(2u - 1) * 180/pi, phi2 = 360 u, written ``%.6f``;
* all random numbers from SplitMix64 (seeds first, then Euler triples).

The reference ships no mesh generator (its testf.c:97-130 only writes
lamellae), so this module is what every synthetic case is built from.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

_MASK = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014), 53-bit uniform doubles."""

    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))


def voronoi_grains(dims, n_grains: int, seed: int):
    """Return (grain_id[V] 1-based int32, euler_deg[n_grains,3]) for a
    periodic Voronoi tessellation of an nx*ny*nz grid (x slowest), computed
    by the library's voce_voronoi (include/voce_batch.h)."""
    import ctypes as C
    from . import load_library
    lib = load_library()
    nx, ny, nz = (int(d) for d in dims)
    gid = np.empty(nx * ny * nz, dtype=np.int32)
    eul = np.empty((n_grains, 3), dtype=np.float64)
    d3 = (C.c_int * 3)(nx, ny, nz)
    rc = lib.voce_voronoi(d3, int(n_grains), C.c_uint64(seed & _MASK), gid.ctypes.data_as(C.POINTER(C.c_int)),
                          eul.ctypes.data_as(C.POINTER(C.c_double)))
    if rc:
        raise RuntimeError("voce_voronoi failed: %d" % rc)
    return gid, eul


def write_mesh(path, dims, grain, euler_by_grain, phase=None):
    """8-column mesh: phi1 Phi phi2 (deg) i j k (1-based) grain phase,
    i slowest (init.c:405-423)."""
    nx, ny, nz = dims
    V = nx * ny * nz
    if phase is None:
        phase = np.ones(V, dtype=np.int32)
    idx = np.arange(V)
    ii = idx // (ny * nz) + 1
    jj = (idx // nz) % ny + 1
    kk = idx % nz + 1
    e = euler_by_grain[grain - 1]
    with open(path, "w") as f:
        for v in range(V):
            f.write("%.6f %.6f %.6f %d %d %d %d %d\n" % (e[v, 0], e[v, 1], e[v, 2], ii[v], jj[v], kk[v],
                                                        grain[v], phase[v]))


FCC_SLIP = [  # Cu_p.sx:24-35, (n) (b)
    ((1, 1, -1), (0, 1, 1)), ((1, 1, -1), (1, 0, 1)), ((1, 1, -1), (1, -1, 0)),
    ((1, -1, -1), (0, 1, -1)), ((1, -1, -1), (1, 0, 1)), ((1, -1, -1), (1, 1, 0)),
    ((1, -1, 1), (0, 1, 1)), ((1, -1, 1), (1, 0, -1)), ((1, -1, 1), (1, 1, 0)),
    ((1, 1, 1), (0, 1, -1)), ((1, 1, 1), (1, 0, -1)), ((1, 1, 1), (1, -1, 0)),
]


@dataclass
class Material:
    nsrs: float = 20.0
    gam0: float = 1.0
    tau0f: float = 16.0
    tau0b: float = 16.0
    tau1: float = 2.0
    theta0: float = 250.0
    theta1: float = 50.0
    slip: list = field(default_factory=lambda: list(FCC_SLIP))


def write_material(path, m: Material):
    """One-mode file in the positional layout PlasticInit reads (init.c:65-177)."""
    lines = [
        "# Crystal system (XtalSys):", "CUBIC", "",
        "# Crystal axex (XtalAxis):", "1.0 1.0 1.0", "",
        "# Total number of modes listed in the file (N_modes_max):", "1", "",
        "# Number of modes to be used in the calculation (N_modes):", "1", "",
        "# Labels of the modes to be used (iMode):", "1", "",
        "# Mode# 1 -- <110>(111) SLIP",
        "\tnsmx\tnrsx\tgamd0x\ttwshx\tisectwx",
        "\t%d\t\t%r\t\t%r\t\t0.0\t\t0" % (len(m.slip), m.nsrs, m.gam0),
        "\ttau0xf\ttau0xb\ttau1x\tthet0\tthet1",
        "\t%r\t\t%r\t\t%r\t\t%r\t\t%r" % (m.tau0f, m.tau0b, m.tau1, m.theta0, m.theta1),
        "\thselfx\thlatex2", "\t1.0\t\t1.0", "\tSlip (n-b)",
    ]
    for n, b in m.slip:
        lines.append("\t%d %d %d\t\t%d %d %d" % (n + b))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


@dataclass
class Case:
    """One RVE run: grid, grains, load and solver controls."""
    dims: tuple = (16, 16, 16)
    n_grains: int = 8
    seed: int = 1
    time_step: float = 1.0
    n_steps: int = 50
    err: float = 1e-5
    iter_max: int = 100
    velgrad_flag: tuple = (1, 0, 0, 1, 1, 1)
    velgrad: tuple = (1.0e-3, -0.5e-3, -0.5e-3, 0.0, 0.0, 0.0)
    stress_flag: tuple = (0, 1, 1, 0, 0, 0)
    stress: tuple = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    elast: tuple = (168400.0, 121400.0, 75400.0, 0.35)
    material: Material = field(default_factory=Material)
    latent: tuple = (1.0, 1.0, 1.0, 1.0, 1.0, 1.0)  # self coplane cross glissile hirth lomer
    update_flag: int = 1
    hard_flag: int = 1
    print_control: tuple = (0, 1)
    tex_flag: int = 0
    # phase II: gas (Type_phases 1) unless made crystalline with its own
    # constants; grains whose id is a multiple of phase2_every get phase 2
    type_phases: tuple = (0, 1)
    elast2: tuple = (0.0, 0.0, 0.0, 0.0)
    material2: Material = None
    phase2_every: int = 0

    @property
    def nvox(self):
        return self.dims[0] * self.dims[1] * self.dims[2]


def _fmt(v):
    if isinstance(v, (float, np.floating)):
        return repr(float(v))
    if isinstance(v, np.integer):
        return str(int(v))
    return str(v)


def write_namelist(path, c: Case, mesh_name="mesh.txt", sx_name="material.sx"):
    """Namelist with every key of io.c:26-51; lower-case ``hirth`` and a
    trailing newline so that no key is dropped."""
    kv = [
        ("N_phases", "2"),
        ("CellDim", " ".join(map(str, c.dims))),
        ("initial_ms", mesh_name),
        ("Type_phases", " ".join(map(str, c.type_phases))),
        ("ElastConst_PhaseI", " ".join(_fmt(float(x)) for x in c.elast)),
        ("Slip_PhaseI", sx_name),
        ("ElastConst_PhaseII", " ".join(_fmt(float(x)) for x in c.elast2)),
        ("Slip_PhaseII", "material2.sx" if c.type_phases[1] == 0 else "XXXXX"),
        ("VelGrad_BC_Flag", " ".join(map(str, c.velgrad_flag))),
        ("VelGrad_BC", " ".join(_fmt(float(x)) for x in c.velgrad)),
        ("Stress_BC_Flag", " ".join(map(str, c.stress_flag))),
        ("Stress_BC", " ".join(_fmt(float(x)) for x in c.stress)),
        ("TimeStep", _fmt(float(c.time_step))),
        ("N_steps", str(c.n_steps)),
        ("Err", _fmt(float(c.err))),
        ("IterMax", str(c.iter_max)),
        ("PrintControl", " ".join(map(str, c.print_control))),
        ("Update_Flag", str(c.update_flag)),
        ("Hard_Flag", str(c.hard_flag)),
        ("Tex_Flag", str(c.tex_flag)),
    ]
    names = ("self", "coplane", "cross", "glissile", "hirth", "lomer")
    for n, v in zip(names, c.latent):
        kv.append((n, "%s 0.0" % _fmt(float(v))))
    lines = ["%s\t%s" % (k, v) for k, v in kv]
    for ln in lines:
        # GetNameList reads 80-byte records (io.c:64): a longer line would be
        # split and its tail silently lost
        if len(ln) >= 79:
            raise ValueError("namelist line longer than the reference reads (79 chars): %r" % ln)
    with open(path, "w") as f:
        f.write("# synthetic case: %s grid, %d grains, seed %d\n" % ("x".join(map(str, c.dims)),
                                                                     c.n_grains, c.seed))
        for ln in lines:
            f.write(ln + "\n")


def write_case(directory, c: Case, grain=None, euler=None):
    """Write input, material.sx and mesh.txt for case ``c`` into ``directory``;
    returns the namelist path."""
    os.makedirs(directory, exist_ok=True)
    if grain is None:
        grain, euler = voronoi_grains(c.dims, c.n_grains, c.seed)
    write_material(os.path.join(directory, "material.sx"), c.material)
    if c.type_phases[1] == 0:
        write_material(os.path.join(directory, "material2.sx"), c.material2 or c.material)
    phase = None
    if c.phase2_every > 0:
        phase = np.where(np.asarray(grain) % c.phase2_every == 0, 2, 1).astype(np.int32)
    write_mesh(os.path.join(directory, "mesh.txt"), c.dims, grain, euler, phase)
    p = os.path.join(directory, "input")
    write_namelist(p, c)
    return p


def config_case(n: int) -> Case:
    """The workloads the metric is quoted on (16^3 CPU case; 64^3 single RVE;
    128^3 plane-strain compression)."""
    if n == 1:
        return Case(dims=(16, 16, 16), n_grains=8, seed=1, time_step=1.0, n_steps=50)
    if n == 2:
        return Case(dims=(64, 64, 64), n_grains=125, seed=2, time_step=0.5, n_steps=100)
    if n == 3:
        return Case(dims=(128, 128, 128), n_grains=1000, seed=3, time_step=1.0, n_steps=100,
                    velgrad_flag=(1, 1, 1, 1, 1, 1), velgrad=(1.0e-3, 0.0, -1.0e-3, 0.0, 0.0, 0.0),
                    stress_flag=(0, 0, 0, 0, 0, 0))
    raise ValueError(n)


# ---------------------------------------------------------------------------
# Many-RVE workloads
# ---------------------------------------------------------------------------

def _traceless_sym(rng):
    a = [2.0 * rng.uniform() - 1.0 for _ in range(6)]
    D = np.array([[a[0], a[3], a[4]], [a[3], a[1], a[5]], [a[4], a[5], a[2]]])
    D -= np.trace(D) / 3.0 * np.eye(3)
    return D


def _skew(rng):
    w = [2.0 * rng.uniform() - 1.0 for _ in range(3)]
    return np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])


def _vm(T):
    return math.sqrt(2.0 / 3.0 * float((T * T).sum()))


@dataclass
class RveSpec:
    """One RVE of a many-RVE workload: grid, grains, seed and velocity gradient."""
    dims: tuple
    n_grains: int
    seed: int
    velgrad: np.ndarray     # 3x3, 1/s


def batch_specs(n=256, sizes=(16, 32, 48, 64), grains=(8, 125), seed0=4000, rate=1e-3, skew=True):
    """Batched-RVE workload: grid N^3 with N uniform over `sizes`, grain count
    uniform over grains[0]..grains[1], Voronoi seed seed0 + r, and
    L = D + W with D random symmetric traceless scaled to von Mises rate
    `rate` and W random skew on the same scale (W = 0 with skew=False)."""
    out = []
    for r in range(n):
        rng = SplitMix64(seed0 + r)
        N = sizes[min(len(sizes) - 1, int(rng.uniform() * len(sizes)))]
        g = grains[0] + min(grains[1] - grains[0], int(rng.uniform() * (grains[1] - grains[0] + 1)))
        D = _traceless_sym(rng)
        D *= rate / _vm(D)
        L = D.copy()
        W = _skew(rng)
        if skew:
            L += W * (rate / _vm(W))
        out.append(RveSpec((N, N, N), g, seed0 + r, L))
    return out


def serve_specs(n_ids=4096, sizes=(16, 32, 64), grains=(8, 64), seed0=5_000_000):
    """Served workload ids: grid, grains and Voronoi seed per id (the velocity
    gradient comes with each request, see request_velgrad)."""
    out = []
    for i in range(n_ids):
        rng = SplitMix64(seed0 + i)
        N = sizes[min(len(sizes) - 1, int(rng.uniform() * len(sizes)))]
        g = grains[0] + min(grains[1] - grains[0], int(rng.uniform() * (grains[1] - grains[0] + 1)))
        out.append(RveSpec((N, N, N), g, seed0 + i, np.zeros((3, 3))))
    return out


def request_velgrad(rng, rate=1e-3):
    """Random traceless velocity gradient with Frobenius norm `rate`."""
    L = np.array([[2.0 * rng.uniform() - 1.0 for _ in range(3)] for _ in range(3)])
    L -= np.trace(L) / 3.0 * np.eye(3)
    return L * (rate / math.sqrt(float((L * L).sum())))


def full_velocity_case(dims=(16, 16, 16), time_step=1.0, err=1e-5, iter_max=100, n_steps=20):
    """A Case whose namelist imposes every velocity-gradient component (the
    per-RVE gradient replaces VelGrad_BC when the case is loaded)."""
    return Case(dims=dims, time_step=time_step, err=err, iter_max=iter_max, n_steps=n_steps,
                velgrad_flag=(1, 1, 1, 1, 1, 1), velgrad=(1.0e-3, -0.5e-3, -0.5e-3, 0.0, 0.0, 0.0),
                stress_flag=(0, 0, 0, 0, 0, 0))


def load_specs(directory, specs, base_case=None):
    """Write one namelist + material into `directory` and load every spec
    with its generated mesh and its own velocity gradient."""
    from . import load_case_voronoi
    base_case = base_case or full_velocity_case()
    write_case(directory, base_case, grain=np.ones(base_case.nvox, dtype=np.int32),
               euler=np.zeros((1, 3)))
    path = os.path.join(directory, "input")
    cases = []
    for sp in specs:
        cs = load_case_voronoi(path, directory, sp.dims, sp.n_grains, sp.seed)
        cs.set_velgrad_full(sp.velgrad)
        cases.append(cs)
    return cases
