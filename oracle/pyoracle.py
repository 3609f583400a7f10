"""ctypes view of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
MAXIT = 1000


class OrcReport(C.Structure):
    _fields_ = [("istep", C.c_int), ("iter", C.c_int), ("iter_total", C.c_int),
                ("e_vm", C.c_double), ("evmp", C.c_double), ("d_vm", C.c_double),
                ("dvmp", C.c_double), ("s_vm", C.c_double), ("s_vm1", C.c_double),
                ("SigAvg", C.c_double * 9), ("SigAvg1", C.c_double * 9),
                ("DisGradAvg_actual", C.c_double * 9), ("VelGradAvg", C.c_double * 9),
                ("EpsAvg", C.c_double * 9), ("EdotAvg", C.c_double * 9),
                ("err_e", C.c_double * MAXIT), ("err_s", C.c_double * MAXIT),
                ("svm_it", C.c_double * MAXIT)]


_lib = None
D = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle not built (sh oracle/build.sh)")
        L = C.CDLL(LIB)
        L.orc_setup.restype = C.c_void_p
        L.orc_setup.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_int)]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_nvox.argtypes = [C.c_void_p]
        L.orc_increment.argtypes = [C.c_void_p, C.POINTER(OrcReport)]
        L.orc_get_field.argtypes = [C.c_void_p, C.c_int, D]
        L.orc_get_macro.argtypes = [C.c_void_p, D]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_write_outputs.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int]
        L.orc_fft3d.argtypes = [D, D, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_kvector.argtypes = [C.c_int] * 6 + [D]
        L.orc_green.argtypes = [D, D, D]
        L.orc_chg_basis.argtypes = [D, D, D, D, C.c_int]
        L.orc_chg_basis5.argtypes = [D, D, D, D, C.c_int]
        L.orc_lu_inv66.argtypes = [D]
        L.orc_euler_to_mat.argtypes = [D, D, D, D, C.c_int]
        L.orc_ten4th_transform.argtypes = [D, D, D, C.c_int]
        L.orc_orientation.argtypes = [D, D]
        L.orc_strain_rate_eval.argtypes = [D, D, D, D, C.c_int, D, D, D, D, D, D]
        L.orc_update_stress_voxel.argtypes = [D, D, D, D, D, D, D, C.c_int, D, D, C.c_double, D, D, D, D,
                                              D, D, D, D]
        L.orc_harden_voxel.argtypes = [D, D, D, D, C.c_int, D, D, D, C.c_double]
        L.orc_hard_matrix.argtypes = [C.c_double] * 6 + [D]
        L.orc_parse_namelist.argtypes = [C.c_char_p, C.POINTER(C.c_int), D, C.c_int]
        L.orc_update_orient_voxel.argtypes = [D, D, D, D, D, D, C.c_int, C.c_double]
        L.orc_update_schmid_voxel.argtypes = [D, D, C.c_int, D]
        IP = C.POINTER(C.c_int)
        L.orc_bc_correction.argtypes = [D, IP, IP, D, D, D]
        L.orc_err_normalise.argtypes = [C.c_double] * 5 + [D]
        L.orc_rho_work_voxel.argtypes = [D, D, D, D, C.c_int, D]
        _lib = L
    return _lib


def ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(D)


FIELD_NCOMP = {0: 6, 1: 6, 2: 6, 3: 6, 4: 9, 5: 12, 6: 12, 7: 12, 8: 12, 9: 12, 10: 9, 11: 1, 12: 1,
               13: 1, 14: 1, 15: 1, 16: 1, 17: 9, 18: 60, 19: 1, 20: 1, 21: 3, 22: 1, 23: 1}
FIELD_IDS = {"sig": 0, "eps": 1, "edot": 2, "els": 3, "disgrad": 4, "crss": 5, "crss_bwd": 6,
             "gamtot": 7, "gamdot": 8, "taux": 9, "tranmat": 10, "gamacum": 11, "svm": 12, "evm": 13,
             "work_p": 14, "rho_rss": 15, "rho_crss": 16, "velgrad": 17, "schmid": 18,
             "nr_iters": 19, "work_e": 20, "new_position": 21, "err_e_v": 22, "err_s_v": 23}


def report_dict(r: OrcReport) -> dict:
    n = r.iter
    d = {k: getattr(r, k) for k in ("istep", "iter", "iter_total", "e_vm", "evmp", "d_vm", "dvmp", "s_vm", "s_vm1")}
    d["sig_avg"] = np.array(r.SigAvg[:]).reshape(3, 3)
    d["sig_avg1"] = np.array(r.SigAvg1[:]).reshape(3, 3)
    d["disgrad_avg"] = np.array(r.DisGradAvg_actual[:]).reshape(3, 3)
    d["velgrad_avg"] = np.array(r.VelGradAvg[:]).reshape(3, 3)
    d["eps_avg"] = np.array(r.EpsAvg[:]).reshape(3, 3)
    d["edot_avg"] = np.array(r.EdotAvg[:]).reshape(3, 3)
    d["err_e"] = np.array(r.err_e[:n])
    d["err_s"] = np.array(r.err_s[:n])
    d["svm_it"] = np.array(r.svm_it[:n])
    return d


class Job:
    """One oracle run over a reference-format input."""

    def __init__(self, input_path: str, base_dir: str | None = None, threads: int | None = None):
        L = lib()
        if threads:
            L.orc_set_threads(int(threads))
        err = C.c_int(0)
        h = L.orc_setup(input_path.encode(), base_dir.encode() if base_dir else None, C.byref(err))
        if not h:
            raise RuntimeError("oracle setup failed: code %d" % err.value)
        self.code = 0
        self._h = h
        self.nvox = L.orc_nvox(h)

    def increment(self) -> dict:
        r = OrcReport()
        rc = lib().orc_increment(self._h, C.byref(r))
        if rc:
            raise RuntimeError("oracle increment failed: code %d" % rc)
        return report_dict(r)

    def field(self, name) -> np.ndarray:
        fid = FIELD_IDS[name] if isinstance(name, str) else int(name)
        out = np.empty((FIELD_NCOMP[fid], self.nvox))
        lib().orc_get_field(self._h, fid, ptr(out))
        return out

    def macro(self) -> np.ndarray:
        out = np.zeros(102)
        lib().orc_get_macro(self._h, ptr(out))
        return out

    def write_outputs(self, directory, step, tex_flag):
        return lib().orc_write_outputs(self._h, directory.encode(), step, tex_flag)

    def close(self):
        if getattr(self, "_h", None):
            lib().orc_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def setup_error(input_path: str, base_dir: str | None = None) -> int:
    """Return the abort code the oracle's setup stops with (0 if it succeeds)."""
    err = C.c_int(0)
    h = lib().orc_setup(input_path.encode(), base_dir.encode() if base_dir else None, C.byref(err))
    if h:
        lib().orc_free(h)
        return 0
    return err.value


def fft3d(re: np.ndarray, im: np.ndarray, sign: int):
    re = np.ascontiguousarray(re, dtype=np.float64).copy()
    im = np.ascontiguousarray(im, dtype=np.float64).copy()
    nx, ny, nz = re.shape
    lib().orc_fft3d(ptr(re), ptr(im), nx, ny, nz, sign)
    return re, im
