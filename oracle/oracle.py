"""ctypes loader for the CPU oracle (libsrx_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libsrx_oracle.so")


class Options(C.Structure):
    _fields_ = [
        ("precision", C.c_int), ("fixed_iters", C.c_int), ("err_override", C.c_double),
        ("itermax_override", C.c_int), ("nsteps_override", C.c_int), ("tau0_mode", C.c_int),
        ("tau0_value", C.c_double), ("svm_mode", C.c_int), ("svm_stray", C.c_double),
        ("write_files", C.c_int), ("write_fields", C.c_int), ("nthreads", C.c_int),
        ("verbose", C.c_int), ("warm_start", C.c_int), ("gnd", C.c_int), ("reserved", C.c_int * 7),
    ]


class Report(C.Structure):
    _fields_ = [
        ("istep", C.c_int), ("iters", C.c_int), ("converged", C.c_int), ("nr_iters_max", C.c_int),
        ("nr_iters_total", C.c_long), ("err_s", C.c_double), ("err_e", C.c_double),
        ("time_tot", C.c_double), ("e_vm", C.c_double), ("evmp", C.c_double), ("d_vm", C.c_double),
        ("dvmp", C.c_double), ("s_vm", C.c_double), ("s_vm1", C.c_double),
        ("rho_F_avg", C.c_double), ("rho_P_avg", C.c_double), ("rho_s_avg", C.c_double),
        ("rho_m_avg", C.c_double), ("rho_g_avg", C.c_double), ("rho_dipole_avg", C.c_double),
        ("rho_forest_avg", C.c_double), ("rho_inplane_avg", C.c_double),
        ("sig_avg", C.c_double * 9), ("sig_avg1", C.c_double * 9), ("eps_avg", C.c_double * 9),
        ("edot_avg", C.c_double * 9), ("disgrad_avg_actual", C.c_double * 9),
        ("velgrad_avg", C.c_double * 9), ("nr_wave_iters", C.c_double),
        ("reserved", C.c_double * 15),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


FIELDS = {"sig": (0, 9), "eps": (1, 9), "edot": (2, 9), "disgrad": (3, 9), "velgrad": (4, 9),
          "tranmat": (5, 9), "rho_s": (6, 12), "rho_m": (7, 12), "rho_F": (8, 12), "rho_P": (9, 12),
          "gamdot": (10, 12), "gam_acum": (11, 12), "rho_dipole": (12, 12), "rho_forest": (13, 12),
          "rho_inplane": (14, 12), "nr_iters": (15, 1), "fluct": (16, 9), "trial_gam_acum": (17, 12),
          "rho_tot": (18, 1), "phase": (19, 1), "grain": (20, 1), "rho_g1": (21, 12), "rho_g2": (22, 12)}

_lib = None


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.POINTER
        L.orc_default_options.argtypes = [P(Options)]
        L.orc_open.argtypes = [C.c_char_p, C.c_char_p, P(Options), P(C.c_void_p)]
        L.orc_step.argtypes = [C.c_void_p, P(Report)]
        L.orc_run.argtypes = [C.c_void_p]
        L.orc_get_field.argtypes = [C.c_void_p, C.c_int, P(C.c_double), C.c_long]
        L.orc_set_field.argtypes = [C.c_void_p, C.c_int, P(C.c_double), C.c_long]
        L.orc_get_setup.argtypes = [C.c_void_p, C.c_int, P(C.c_double), C.c_long]
        L.orc_dims.argtypes = [C.c_void_p, P(C.c_int)]
        L.orc_close.argtypes = [C.c_void_p]
        L.orc_chg_basis.argtypes = [P(C.c_double)] * 4 + [C.c_int, C.c_int]
        L.orc_lu_inv66.argtypes = [P(C.c_double)]
        L.orc_euler_to_matrix.argtypes = [P(C.c_double)] * 4 + [C.c_int]
        L.orc_fft3d.argtypes = [P(C.c_double), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_green_point.argtypes = [P(C.c_double), P(C.c_double), C.c_int, C.c_int, P(C.c_double),
                                      P(C.c_double), P(C.c_double)]
        L.orc_strain_rate.argtypes = [C.c_void_p, C.c_int, P(C.c_double), P(C.c_double), P(C.c_double)]
        L.orc_nr_voxel.argtypes = [C.c_void_p, C.c_int, P(C.c_double), P(C.c_double)]
        L.orc_update_schmid.argtypes = [C.c_void_p]
        L.orc_get_schmid.argtypes = [C.c_void_p, C.c_int, P(C.c_double)]
        L.orc_orientation.argtypes = [P(C.c_double), P(C.c_double)]
        L.orc_update_orient.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def options(**kw) -> Options:
    o = Options()
    lib().orc_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Oracle:
    """One oracle run: orc_open(deck) ... orc_step() ... orc_close()."""

    def __init__(self, deck: str, workdir: str | None = None, **opt):
        self.L = lib()
        self.h = C.c_void_p()
        self.opt = options(**opt)
        rc = self.L.orc_open(deck.encode(), (workdir or os.path.dirname(os.path.abspath(deck))).encode(),
                             C.byref(self.opt), C.byref(self.h))
        if rc:
            raise RuntimeError("orc_open failed with code %d" % rc)
        n = (C.c_int * 3)()
        self.L.orc_dims(self.h, n)
        self.n = tuple(n)
        self.nxyz = n[0] * n[1] * n[2]

    def step(self) -> dict:
        r = Report()
        rc = self.L.orc_step(self.h, C.byref(r))
        if rc:
            raise RuntimeError("orc_step failed with code %d" % rc)
        return r.as_dict()

    def field(self, name: str) -> np.ndarray:
        fid, nc = FIELDS[name]
        out = np.zeros(self.nxyz * nc)
        rc = self.L.orc_get_field(self.h, fid, dptr(out), out.size)
        if rc:
            raise RuntimeError("orc_get_field rc=%d" % rc)
        return out.reshape(self.nxyz, nc) if nc > 1 else out

    def set_field(self, name: str, a: np.ndarray) -> None:
        fid, nc = FIELDS[name]
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        rc = self.L.orc_set_field(self.h, fid, dptr(a), a.size)
        if rc:
            raise RuntimeError("orc_set_field rc=%d" % rc)

    def setup(self, which: int, n: int) -> np.ndarray:
        out = np.zeros(n)
        rc = self.L.orc_get_setup(self.h, which, dptr(out), n)
        if rc:
            raise RuntimeError("orc_get_setup rc=%d" % rc)
        return out

    def close(self):
        if self.h:
            self.L.orc_close(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def open_status(deck: str, workdir: str | None = None, **opt) -> int:
    """Return the integer code orc_open gives for a deck (0 on success)."""
    L = lib()
    h = C.c_void_p()
    o = options(**opt)
    rc = L.orc_open(deck.encode(), (workdir or os.path.dirname(os.path.abspath(deck))).encode(), C.byref(o), C.byref(h))
    if rc == 0:
        L.orc_close(h)
    return rc
