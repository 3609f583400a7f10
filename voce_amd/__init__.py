"""voce_amd -- MI355X EVP-FFT crystal-plasticity solver (Python view of the C ABI).

The solver itself is the native library ``voce_amd/libvoce.so`` (HIP kernels
for gfx950 behind the C ABI of ``include/voce.h``).  This module binds that
ABI with ctypes; it holds no solver logic and has no fallback: if the
library is missing, every entry point raises.

The Python names follow the reference's boundary (SetupJob / Evolution /
DestroyJob, init.c:617-930 and evolution.c:265-622):

    case = load_case("input")               # GetNameList + .sx + mesh
    s = Solver(case.setup, precision="fp64")  # SetupJob
    rep = s.increment()                     # one istep of Evolution
    sig = s.field("sig")                    # (6, V) like WriteSigMPI's data
    s.close()                               # DestroyJob
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvoce.so")

NSYS = 12
FP64, MIXED = 0, 1

FIELDS = {
    "sig": 0, "eps": 1, "edot": 2, "els": 3, "disgrad": 4, "crss": 5, "crss_bwd": 6,
    "gamtot": 7, "gamdot": 8, "taux": 9, "tranmat": 10, "gamacum": 11, "svm": 12, "evm": 13,
    "work_p": 14, "rho_rss": 15, "rho_crss": 16, "velgrad": 17, "schmid": 18, "work_e": 19,
    "new_position": 20,
}


class VoceError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        msg = "voce error %d" % code
        if _lib is not None:
            msg += ": " + _lib.voce_strerror(code).decode()
        if what:
            msg = what + ": " + msg
        super().__init__(msg)


class Phase(C.Structure):
    _fields_ = [("type", C.c_int), ("elast", C.c_double * 3), ("nsys", C.c_int),
                ("bn", (C.c_double * 3) * NSYS), ("bb", (C.c_double * 3) * NSYS),
                ("nsrs", C.c_double), ("gam0", C.c_double), ("tau0f", C.c_double),
                ("tau0b", C.c_double), ("tau1", C.c_double), ("thet0", C.c_double),
                ("thet1", C.c_double), ("hard", (C.c_double * NSYS) * NSYS)]


class Setup(C.Structure):
    _fields_ = [("dims", C.c_int * 3), ("n_phases", C.c_int), ("phase", Phase * 2),
                ("time_step", C.c_double), ("err", C.c_double), ("iter_max", C.c_int),
                ("update_flag", C.c_int), ("hard_flag", C.c_int),
                ("velgrad_flag", C.c_int * 6), ("velgrad", C.c_double * 6),
                ("stress_flag", C.c_int * 6), ("stress", C.c_double * 6),
                ("use_full_velgrad", C.c_int), ("velgrad_full", C.c_double * 9),
                ("euler_deg", C.POINTER(C.c_double)), ("grain", C.POINTER(C.c_int)),
                ("phase_id", C.POINTER(C.c_int)), ("precision", C.c_int)]


class Report(C.Structure):
    _fields_ = [("istep", C.c_int), ("iter", C.c_int), ("iter_total", C.c_int),
                ("e_vm", C.c_double), ("evmp", C.c_double), ("d_vm", C.c_double),
                ("dvmp", C.c_double), ("s_vm", C.c_double), ("s_vm1", C.c_double),
                ("sig_avg", C.c_double * 9), ("sig_avg1", C.c_double * 9),
                ("disgrad_avg", C.c_double * 9), ("velgrad_avg", C.c_double * 9),
                ("eps_avg", C.c_double * 9), ("edot_avg", C.c_double * 9),
                ("err_e", C.POINTER(C.c_double)), ("err_s", C.POINTER(C.c_double)),
                ("svm_it", C.POINTER(C.c_double))]


class Case(C.Structure):
    _fields_ = [("setup", Setup), ("n_steps", C.c_int), ("print_control", C.c_int * 2),
                ("tex_flag", C.c_int), ("namelist_status", C.c_int * 26),
                ("euler_deg", C.POINTER(C.c_double)), ("grain", C.POINTER(C.c_int)),
                ("phase_id", C.POINTER(C.c_int))]


# every symbol include/voce.h declares
EXPORTS = ["voce_field_ncomp", "voce_ctx_create", "voce_increment", "voce_increment_noalloc", "voce_get_field",
           "voce_set_diagnostics", "voce_ctx_nvox", "voce_ctx_destroy", "voce_strerror",
           "voce_load_case", "voce_case_free", "voce_write_fields",
           # include/voce_prof.h
           "voce_prof_enable", "voce_prof_read", "voce_ctx_info",
           # include/voce_batch.h
           "voce_autotune", "voce_apply_tuning", "voce_get_tuning", "voce_batch_create", "voce_batch_increment",
           "voce_batch_size", "voce_batch_member", "voce_batch_info", "voce_batch_destroy",
           "voce_set_velgrad", "voce_voronoi", "voce_load_case_voronoi",
           "voce_server_create", "voce_server_add_rve", "voce_server_submit", "voce_server_step",
           "voce_server_poll", "voce_server_info", "voce_server_rve", "voce_server_destroy",
           "voce_server_inject_fault"]

PROF_KERNELS = ["plane_fwd", "xgreen", "plane_inv", "nr", "fin_iter"]

_lib = None


def load_library(path: str = LIB_PATH):
    """Load libvoce.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise VoceError(-6, "native library %s not built (python -m voce_amd.build)" % path)
    lib = C.CDLL(path)
    lib.voce_strerror.restype = C.c_char_p
    lib.voce_strerror.argtypes = [C.c_int]
    lib.voce_field_ncomp.argtypes = [C.c_int]
    lib.voce_ctx_create.argtypes = [C.POINTER(Setup), C.POINTER(C.c_void_p)]
    lib.voce_increment.argtypes = [C.c_void_p, C.POINTER(Report)]
    lib.voce_increment_noalloc.argtypes = [C.c_void_p, C.POINTER(Report)]
    lib.voce_get_field.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_size_t]
    lib.voce_set_diagnostics.argtypes = [C.c_void_p, C.c_int]
    lib.voce_ctx_nvox.restype = C.c_long
    lib.voce_ctx_nvox.argtypes = [C.c_void_p]
    lib.voce_ctx_destroy.argtypes = [C.c_void_p]
    lib.voce_ctx_destroy.restype = None
    lib.voce_load_case.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(Case)]
    lib.voce_case_free.argtypes = [C.POINTER(Case)]
    lib.voce_case_free.restype = None
    lib.voce_write_fields.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_int]
    lib.voce_prof_enable.argtypes = [C.c_void_p, C.c_int]
    lib.voce_prof_read.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_long)]
    lib.voce_ctx_info.argtypes = [C.c_void_p, C.POINTER(C.c_long)]
    lib.voce_voronoi.argtypes = [C.POINTER(C.c_int), C.c_int, C.c_uint64, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    lib.voce_load_case_voronoi.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_int), C.c_int, C.c_uint64,
                                           C.POINTER(Case)]
    lib.voce_autotune.argtypes = [C.c_void_p, C.POINTER(TuneChoice)]
    lib.voce_apply_tuning.argtypes = [C.c_void_p, C.POINTER(TuneChoice)]
    lib.voce_get_tuning.argtypes = [C.c_void_p, C.POINTER(TuneChoice)]
    lib.voce_set_velgrad.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    lib.voce_batch_create.argtypes = [C.POINTER(Setup), C.c_int, C.c_int, C.POINTER(C.c_void_p)]
    lib.voce_batch_increment.argtypes = [C.c_void_p, C.POINTER(Report)]
    lib.voce_batch_size.argtypes = [C.c_void_p]
    lib.voce_batch_member.argtypes = [C.c_void_p, C.c_int]
    lib.voce_batch_member.restype = C.c_void_p
    lib.voce_batch_info.argtypes = [C.c_void_p, C.POINTER(C.c_long)]
    lib.voce_batch_destroy.argtypes = [C.c_void_p]
    lib.voce_batch_destroy.restype = None
    lib.voce_server_create.argtypes = [C.POINTER(ServerOpts), C.POINTER(C.c_void_p)]
    lib.voce_server_add_rve.argtypes = [C.c_void_p, C.c_int, C.POINTER(Setup)]
    lib.voce_server_submit.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_int, C.c_long]
    lib.voce_server_step.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    lib.voce_server_poll.argtypes = [C.c_void_p, C.POINTER(Result), C.c_int]
    lib.voce_server_info.argtypes = [C.c_void_p, C.POINTER(C.c_long)]
    lib.voce_server_rve.argtypes = [C.c_void_p, C.c_int]
    lib.voce_server_rve.restype = C.c_void_p
    lib.voce_server_inject_fault.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.voce_server_destroy.argtypes = [C.c_void_p]
    lib.voce_server_destroy.restype = None
    _lib = lib
    return lib


class ServerOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("max_rves", C.c_int), ("max_inflight", C.c_int),
                ("workspace_bytes", C.c_size_t), ("tune", C.c_int), ("chunk", C.c_int)]


class Result(C.Structure):
    _fields_ = [("tag", C.c_long), ("id", C.c_int), ("status", C.c_int), ("audited", C.c_int),
                ("rep", Report), ("audit_sig_rel", C.c_double),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]


class TuneChoice(C.Structure):
    _fields_ = [("plv", C.c_int), ("xv", C.c_int), ("tk", C.c_int), ("pl_r8", C.c_int), ("x_r8", C.c_int),
                ("us_plane", C.c_double), ("us_x", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LoadedCase:
    """A parsed input (namelist + .sx + mesh); owns the C arrays.  With
    `voronoi=(dims, n_grains, seed)` the mesh is generated instead of read."""

    def __init__(self, path: str, base_dir: str | None = None, voronoi=None):
        lib = load_library()
        self._c = Case()
        b = base_dir.encode() if base_dir else None
        if voronoi is None:
            rc = lib.voce_load_case(path.encode(), b, C.byref(self._c))
        else:
            dims, n, seed = voronoi
            d3 = (C.c_int * 3)(*[int(x) for x in dims])
            rc = lib.voce_load_case_voronoi(path.encode(), b, d3, int(n), C.c_uint64(int(seed) & ((1 << 64) - 1)),
                                            C.byref(self._c))
        if rc:
            raise VoceError(rc, "voce_load_case(%s)" % path)
        self._freed = False

    def set_velgrad_full(self, L):
        """Full-velocity boundary conditions with a general (non-symmetric)
        macroscopic velocity gradient L (3x3, 1/s)."""
        s = self._c.setup
        s.use_full_velgrad = 1
        for i, x in enumerate(np.asarray(L, dtype=np.float64).ravel()):
            s.velgrad_full[i] = float(x)
        for i in range(6):
            s.velgrad_flag[i] = 1
            s.stress_flag[i] = 0

    @property
    def setup(self) -> Setup:
        return self._c.setup

    @property
    def n_steps(self) -> int:
        return self._c.n_steps

    @property
    def namelist_status(self):
        return list(self._c.namelist_status)

    @property
    def nvox(self):
        d = self._c.setup.dims
        return d[0] * d[1] * d[2]

    def arrays(self):
        V = self.nvox
        e = np.ctypeslib.as_array(self._c.euler_deg, shape=(V * 3,)).reshape(V, 3).copy()
        g = np.ctypeslib.as_array(self._c.grain, shape=(V,)).copy()
        p = np.ctypeslib.as_array(self._c.phase_id, shape=(V,)).copy()
        return e, g, p

    def free(self):
        if not self._freed:
            load_library().voce_case_free(C.byref(self._c))
            self._freed = True

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def load_case(path: str, base_dir: str | None = None) -> LoadedCase:
    return LoadedCase(path, base_dir)


def load_case_voronoi(path: str, base_dir, dims, n_grains: int, seed: int) -> LoadedCase:
    """Namelist and materials from `path`, mesh from voce_voronoi."""
    return LoadedCase(path, base_dir, voronoi=(dims, n_grains, seed))


def report_dict(r: Report) -> dict:
    n = r.iter
    d = {k: getattr(r, k) for k in ("istep", "iter", "iter_total", "e_vm", "evmp", "d_vm", "dvmp", "s_vm", "s_vm1")}
    for k in ("sig_avg", "sig_avg1", "disgrad_avg", "velgrad_avg", "eps_avg", "edot_avg"):
        d[k] = np.array(getattr(r, k)[:]).reshape(3, 3)
    # per-iteration logs (NULL in server results)
    d["err_e"] = np.array(r.err_e[:n]) if n and r.err_e else np.zeros(0)
    d["err_s"] = np.array(r.err_s[:n]) if n and r.err_s else np.zeros(0)
    d["svm_it"] = np.array(r.svm_it[:n]) if n and r.svm_it else np.zeros(0)
    return d


class Solver:
    """One RVE resident on the GPU (a voce_ctx)."""

    def __init__(self, setup: Setup, precision="fp64", keepalive=None):
        lib = load_library()
        setup.precision = MIXED if precision in ("mixed", MIXED) else FP64
        self._keep = keepalive
        h = C.c_void_p()
        rc = lib.voce_ctx_create(C.byref(setup), C.byref(h))
        if rc:
            raise VoceError(rc, "voce_ctx_create")
        self._h = h
        self.nvox = lib.voce_ctx_nvox(h)

    @classmethod
    def from_case(cls, case: LoadedCase, precision="fp64"):
        return cls(case.setup, precision, keepalive=case)

    def increment(self) -> dict:
        r = Report()
        rc = load_library().voce_increment(self._h, C.byref(r))
        if rc:
            raise VoceError(rc, "voce_increment")
        return report_dict(r)

    def set_diagnostics(self, on=True):
        rc = load_library().voce_set_diagnostics(self._h, int(bool(on)))
        if rc:
            raise VoceError(rc, "voce_set_diagnostics")

    def field(self, name) -> np.ndarray:
        fid = FIELDS[name] if isinstance(name, str) else int(name)
        lib = load_library()
        nc = lib.voce_field_ncomp(fid)
        out = np.empty((nc, self.nvox), dtype=np.float64)
        rc = lib.voce_get_field(self._h, fid, out.ctypes.data_as(C.POINTER(C.c_double)), out.nbytes)
        if rc:
            raise VoceError(rc, "voce_get_field(%s)" % name)
        return out

    def write_fields(self, directory, step, tex_flag=0, final=False):
        rc = load_library().voce_write_fields(self._h, directory.encode(), step, tex_flag, int(final))
        if rc:
            raise VoceError(rc, "voce_write_fields")

    def autotune(self) -> dict:
        ch = TuneChoice()
        rc = load_library().voce_autotune(self._h, C.byref(ch))
        if rc:
            raise VoceError(rc, "voce_autotune")
        return ch.as_dict()

    def tuning(self) -> dict:
        """The current FFT launch choice (plv, xv, tk, pl_r8, x_r8)."""
        ch = TuneChoice()
        rc = load_library().voce_get_tuning(self._h, C.byref(ch))
        if rc:
            raise VoceError(rc, "voce_get_tuning")
        return {k: v for k, v in ch.as_dict().items() if not k.startswith("us_")}

    def apply_tuning(self, choice: dict):
        ch = TuneChoice(**{k: int(choice[k]) for k in ("plv", "xv", "tk", "pl_r8", "x_r8")})
        rc = load_library().voce_apply_tuning(self._h, C.byref(ch))
        if rc:
            raise VoceError(rc, "voce_apply_tuning")

    def set_velgrad(self, L):
        a = (C.c_double * 9)(*[float(x) for x in np.asarray(L, dtype=np.float64).ravel()])
        rc = load_library().voce_set_velgrad(self._h, a)
        if rc:
            raise VoceError(rc, "voce_set_velgrad")

    def prof_enable(self, on=True, every=1):
        """Kernel timing with HIP events on every `every`-th FFT iteration
        of each increment (off when `on` is false)."""
        load_library().voce_prof_enable(self._h, max(1, int(every)) if on else 0)

    def prof_read(self) -> dict:
        ms = (C.c_double * 5)()
        n = (C.c_long * 5)()
        load_library().voce_prof_read(self._h, ms, n)
        return {k: (ms[i], n[i]) for i, k in enumerate(PROF_KERNELS)}

    def close(self):
        if getattr(self, "_h", None):
            if not getattr(self, "_borrowed", False):
                load_library().voce_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Member(Solver):
    """A batch member viewed as a context (owned by its batch)."""

    def __init__(self, handle, nvox, batch):
        self._h = C.c_void_p(handle)
        self._borrowed = True
        self._keep = batch
        self.nvox = nvox


class Batch:
    """Many RVEs advanced together (include/voce_batch.h)."""

    def __init__(self, cases, precision="fp64", tune=True):
        lib = load_library()
        self._cases = list(cases)
        n = len(self._cases)
        arr = (Setup * n)()
        for i, cs in enumerate(self._cases):
            arr[i] = cs.setup
            arr[i].precision = MIXED if precision in ("mixed", MIXED) else FP64
        h = C.c_void_p()
        rc = lib.voce_batch_create(arr, n, int(bool(tune)), C.byref(h))
        if rc:
            raise VoceError(rc, "voce_batch_create")
        self._h = h
        self.n = n
        self.nvox = [cs.nvox for cs in self._cases]

    def increment(self):
        reps = (Report * self.n)()
        rc = load_library().voce_batch_increment(self._h, reps)
        if rc:
            raise VoceError(rc, "voce_batch_increment")
        return [report_dict(r) for r in reps]

    def member(self, i) -> Solver:
        hm = load_library().voce_batch_member(self._h, int(i))
        if not hm:
            raise IndexError(i)
        return _Member(hm, self.nvox[i], self)

    def info(self) -> dict:
        out = (C.c_long * 8)()
        load_library().voce_batch_info(self._h, out)
        keys = ("members", "groups", "voxels", "graph_replays", "graphs", "iterations_launched", "last_max_iter",
                "device_bytes")
        return dict(zip(keys, list(out)))

    def close(self):
        if getattr(self, "_h", None):
            load_library().voce_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """FE^2-style request server over resident RVEs (include/voce_batch.h)."""

    INFO_KEYS = ("registered", "resident_bytes", "pool_bytes", "pool_used", "inflight", "queued", "completed",
                 "audits", "steps", "iterations_launched")

    def __init__(self, max_rves, precision="mixed", max_inflight=0, workspace_bytes=0, tune=True, chunk=16):
        lib = load_library()
        o = ServerOpts(MIXED if precision in ("mixed", MIXED) else FP64, int(max_rves), int(max_inflight),
                       int(workspace_bytes), int(bool(tune)), int(chunk))
        h = C.c_void_p()
        rc = lib.voce_server_create(C.byref(o), C.byref(h))
        if rc:
            raise VoceError(rc, "voce_server_create")
        self._h = h
        self._keep = {}
        self._nvox = {}

    def add_rve(self, rid, case):
        rc = load_library().voce_server_add_rve(self._h, int(rid), C.byref(case.setup))
        if rc:
            raise VoceError(rc, "voce_server_add_rve(%d)" % rid)
        self._nvox[int(rid)] = case.nvox

    def submit(self, rid, L, audit=False, tag=0):
        a = (C.c_double * 9)(*[float(x) for x in np.asarray(L, dtype=np.float64).ravel()])
        rc = load_library().voce_server_submit(self._h, int(rid), a, int(bool(audit)), int(tag))
        if rc:
            raise VoceError(rc, "voce_server_submit")

    def step(self) -> int:
        n = C.c_int(0)
        rc = load_library().voce_server_step(self._h, C.byref(n))
        if rc:
            raise VoceError(rc, "voce_server_step")
        return n.value

    def poll(self, max_n=4096):
        buf = (Result * max_n)()
        n = load_library().voce_server_poll(self._h, buf, max_n)
        out = []
        for r in buf[:n]:
            d = report_dict(r.rep)
            d.update(tag=r.tag, id=r.id, status=r.status, audited=r.audited, audit_sig_rel=r.audit_sig_rel,
                     t_submit=r.t_submit, t_admit=r.t_admit, t_done=r.t_done)
            out.append(d)
        return out

    def drain(self):
        """Step until nothing is queued or in flight; return all results."""
        out = []
        while True:
            n = self.step()
            out += self.poll()
            info = self.info()
            if n == 0 and info["queued"] == 0:
                break
        return out + self.poll()

    def info(self) -> dict:
        out = (C.c_long * 10)()
        load_library().voce_server_info(self._h, out)
        return dict(zip(self.INFO_KEYS, list(out)))

    def inject_fault(self, kind=1, after=1):
        """Failure-handling test hook: the `after`-th next audit-shadow
        allocation reports VOCE_ERR_NOMEM (kind 1, VOCE_FAULT_SHADOW_ALLOC)."""
        rc = load_library().voce_server_inject_fault(self._h, int(kind), int(after))
        if rc:
            raise VoceError(rc, "voce_server_inject_fault")

    def rve(self, rid) -> Solver:
        hm = load_library().voce_server_rve(self._h, int(rid))
        if not hm:
            raise KeyError(rid)
        return _Member(hm, self._nvox[int(rid)], self)

    def close(self):
        if getattr(self, "_h", None):
            load_library().voce_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
