"""srx_amd -- MI355X EVP-FFT crystal-plasticity solver (ctypes front end of libsrx.so).

The product is the C ABI in include/srx.h, implemented by the in-tree
``srx_amd/libsrx.so`` (HIP kernels for gfx950 + C++ host driver).  This module
loads that library and mirrors its entry points for Python callers; it has no
fallback: if the library is missing, importing ``Solver`` raises.

Reference interface mirrored (evp.c:7-31):
    GetNameList + SetupJob  -> Solver(deck)          (srx_open)
    one istep of Evolution  -> Solver.step()         (srx_step)
    Evolution               -> Solver.run()          (srx_run)
    DestroyJob              -> Solver.close()        (srx_close)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsrx.so")

PREC_FP64 = 0
PREC_MIXED = 1
SRX_E_GRID = 2002
SRX_E_SIZE = 2008


class Options(C.Structure):
    _fields_ = [
        ("precision", C.c_int), ("fixed_iters", C.c_int), ("err_override", C.c_double),
        ("itermax_override", C.c_int), ("nsteps_override", C.c_int), ("tau0_mode", C.c_int),
        ("tau0_value", C.c_double), ("svm_mode", C.c_int), ("svm_stray", C.c_double),
        ("write_files", C.c_int), ("write_fields", C.c_int), ("nthreads", C.c_int),
        ("verbose", C.c_int), ("warm_start", C.c_int), ("gnd", C.c_int), ("reserved", C.c_int * 7),
    ]


class StepReport(C.Structure):
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

EXPORTS = ["srx_open", "srx_step", "srx_run", "srx_get_field", "srx_set_field", "srx_dims", "srx_close",
           "srx_strerror", "srx_default_options", "srx_kernel_times", "srx_fft_forward", "srx_fft_inverse",
           "srx_run_ensemble", "srx_mlp_default_train_options", "srx_mlp_train", "srx_mlp_grad",
           "srx_mlp_forward", "srx_mlp_set", "srx_mlp_record", "srx_mlp_samples", "srx_fft_tiles", "srx_hbm_stream",
           "srx_fft_plan", "srx_set_fft_plan", "srx_fft_plan_candidates"]

MLP_BLOB = 6210     # SRX_MLP_BLOB
FFT_MAX_STAGES = 12   # SRX_FFT_MAX_STAGES
MLP_NPARAM = 6150   # SRX_MLP_NPARAM


class MlpTrainOpts(C.Structure):
    """srx_mlp_train_opts of include/srx.h."""
    _fields_ = [("epochs", C.c_int), ("batch", C.c_int), ("lr", C.c_float), ("lr_final", C.c_float),
                ("seed", C.c_ulonglong), ("init", C.c_int), ("reserved", C.c_int * 7)]


class EnsembleResult(C.Structure):
    """srx_ensemble_result of include/srx.h."""
    _fields_ = [("index", C.c_int), ("rc", C.c_int), ("increments", C.c_int), ("voxels", C.c_long),
                ("setup_seconds", C.c_double), ("run_seconds", C.c_double), ("run_phase_seconds", C.c_double),
                ("last", StepReport)]

_lib = None


def lib():
    """Load libsrx.so (built by srx_amd/build.py).  Raises if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libsrx.so is not built (run `python srx_amd/build.py`); "
                               "there is no fallback path")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.srx_default_options.argtypes = [P(Options)]
        L.srx_open.argtypes = [C.c_char_p, C.c_char_p, P(Options), P(C.c_void_p)]
        L.srx_step.argtypes = [C.c_void_p, P(StepReport)]
        L.srx_run.argtypes = [C.c_void_p]
        L.srx_get_field.argtypes = [C.c_void_p, C.c_int, P(C.c_double), C.c_size_t]
        L.srx_set_field.argtypes = [C.c_void_p, C.c_int, P(C.c_double), C.c_size_t]
        L.srx_dims.argtypes = [C.c_void_p, P(C.c_int)]
        L.srx_close.argtypes = [C.c_void_p]
        L.srx_strerror.argtypes = [C.c_int]
        L.srx_strerror.restype = C.c_char_p
        L.srx_kernel_times.argtypes = [C.c_void_p, P(C.c_char_p), P(C.c_double), P(C.c_int), P(C.c_int)]
        L.srx_fft_forward.argtypes = [P(C.c_double), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P(C.c_double)]
        L.srx_fft_inverse.argtypes = [P(C.c_double), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                      P(C.c_double)]
        L.srx_run_ensemble.argtypes = [P(C.c_char_p), P(C.c_char_p), C.c_int, P(Options), C.c_int,
                                       P(EnsembleResult)]
        F = P(C.c_float)
        L.srx_mlp_default_train_options.argtypes = [P(MlpTrainOpts)]
        L.srx_mlp_default_train_options.restype = None
        L.srx_mlp_train.argtypes = [F, F, C.c_size_t, P(MlpTrainOpts), F, F]
        L.srx_mlp_grad.argtypes = [F, F, F, C.c_size_t, F, P(C.c_double)]
        L.srx_mlp_forward.argtypes = [F, F, C.c_size_t, F]
        L.srx_mlp_set.argtypes = [C.c_void_p, F]
        L.srx_mlp_record.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        L.srx_mlp_samples.argtypes = [C.c_void_p, F, F, C.c_size_t, P(C.c_size_t)]
        L.srx_fft_tiles.argtypes = [C.c_void_p, P(C.c_int), P(C.c_double)]
        L.srx_hbm_stream.argtypes = [C.c_size_t, C.c_int, P(C.c_double), P(C.c_double), P(C.c_int), P(C.c_int)]
        L.srx_fft_plan.argtypes = [C.c_void_p, C.c_int, P(C.c_int), P(C.c_int)]
        L.srx_set_fft_plan.argtypes = [C.c_void_p, C.c_int, P(C.c_int), C.c_int]
        L.srx_fft_plan_candidates.argtypes = [C.c_int, P(C.c_int), C.c_int, P(C.c_int)]
        _lib = L
    return _lib


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


class SrxError(RuntimeError):
    def __init__(self, code: int, where: str):
        self.code = code
        msg = lib().srx_strerror(code).decode()
        super().__init__("%s: %s (code %d)" % (where, msg, code))


def fft_plan_candidates(n: int) -> list:
    """Radix plans the FFT autotuner may time for extent n, the default first (host only)."""
    L = lib()
    cnt = C.c_int()
    rc = L.srx_fft_plan_candidates(int(n), None, 0, C.byref(cnt))
    if rc not in (0, SRX_E_SIZE):
        raise SrxError(rc, "srx_fft_plan_candidates")
    buf = (C.c_int * (cnt.value * FFT_MAX_STAGES))()
    rc = L.srx_fft_plan_candidates(int(n), buf, cnt.value, C.byref(cnt))
    if rc:
        raise SrxError(rc, "srx_fft_plan_candidates")
    rows = [list(buf[i * FFT_MAX_STAGES:(i + 1) * FFT_MAX_STAGES]) for i in range(cnt.value)]
    return [[x for x in r if x] for r in rows]


def options(**kw) -> Options:
    o = Options()
    lib().srx_default_options(C.byref(o))
    for k, v in kw.items():
        if k == "reserved":
            for i, x in enumerate(v):
                o.reserved[i] = x
        else:
            setattr(o, k, v)
    return o


class Solver:
    """One solver context (srx_open ... srx_close)."""

    def __init__(self, deck: str, workdir: str | None = None, **opt):
        self.L = lib()
        self.h = C.c_void_p()
        self.opt = options(**opt)
        wd = workdir or os.path.dirname(os.path.abspath(deck))
        rc = self.L.srx_open(deck.encode(), wd.encode(), C.byref(self.opt), C.byref(self.h))
        if rc:
            raise SrxError(rc, "srx_open")
        n = (C.c_int * 3)()
        self.L.srx_dims(self.h, n)
        self.n = tuple(n)
        self.nxyz = n[0] * n[1] * n[2]

    def step(self) -> dict:
        r = StepReport()
        rc = self.L.srx_step(self.h, C.byref(r))
        if rc:
            raise SrxError(rc, "srx_step")
        return r.as_dict()

    def run(self) -> None:
        rc = self.L.srx_run(self.h)
        if rc:
            raise SrxError(rc, "srx_run")

    def field(self, name: str) -> np.ndarray:
        fid, nc = FIELDS[name]
        out = np.zeros(self.nxyz * nc)
        rc = self.L.srx_get_field(self.h, fid, _dptr(out), out.size)
        if rc:
            raise SrxError(rc, "srx_get_field")
        return out.reshape(self.nxyz, nc) if nc > 1 else out

    def set_field(self, name: str, a: np.ndarray) -> None:
        fid, nc = FIELDS[name]
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        rc = self.L.srx_set_field(self.h, fid, _dptr(a), a.size)
        if rc:
            raise SrxError(rc, "srx_set_field")

    def kernel_times(self) -> dict:
        n = C.c_int(16)
        names = (C.c_char_p * 16)()
        ms = (C.c_double * 16)()
        calls = (C.c_int * 16)()
        self.L.srx_kernel_times(self.h, names, ms, calls, C.byref(n))
        return {names[i].decode(): (ms[i], calls[i]) for i in range(n.value)}

    def fft_tiles(self) -> tuple[list, list]:
        """Rows per block chosen by the tile autotuner for the five FFT passes, and their times (us)."""
        t = (C.c_int * 5)()
        us = (C.c_double * 5)()
        rc = self.L.srx_fft_tiles(self.h, t, us)
        if rc:
            raise SrxError(rc, "srx_fft_tiles")
        return list(t), list(us)

    def fft_plan(self, axis: int) -> list:
        """Radices of the FFT plan of one axis (0: x, 1: y, 2: z)."""
        r = (C.c_int * FFT_MAX_STAGES)()
        ns = C.c_int()
        rc = self.L.srx_fft_plan(self.h, int(axis), r, C.byref(ns))
        if rc:
            raise SrxError(rc, "srx_fft_plan")
        return list(r)[:ns.value]

    def set_fft_plan(self, axis: int, radices) -> None:
        """Run the passes of one axis with these radices (product: the extent)."""
        r = (C.c_int * max(1, len(radices)))(*[int(x) for x in radices])
        rc = self.L.srx_set_fft_plan(self.h, int(axis), r, len(radices))
        if rc:
            raise SrxError(rc, "srx_set_fft_plan")

    def mlp_set(self, blob: np.ndarray) -> None:
        """Install the learned Newton start (context opened with warm_start=2)."""
        b = np.ascontiguousarray(blob, dtype=np.float32)
        assert b.size == MLP_BLOB
        rc = self.L.srx_mlp_set(self.h, _fptr(b))
        if rc:
            raise SrxError(rc, "srx_mlp_set")

    def mlp_record(self, stride: int, capacity: int) -> None:
        """Record training samples of voxels 0, stride, ... (stride 0: stop and free)."""
        rc = self.L.srx_mlp_record(self.h, int(stride), int(capacity))
        if rc:
            raise SrxError(rc, "srx_mlp_record")

    def mlp_samples(self) -> tuple[np.ndarray, np.ndarray]:
        """Recorded samples (features [n][24], targets [n][6])."""
        n = C.c_size_t()
        rc = self.L.srx_mlp_samples(self.h, None, None, 0, C.byref(n))
        if rc:
            raise SrxError(rc, "srx_mlp_samples")
        x = np.zeros((n.value, 24), np.float32)
        t = np.zeros((n.value, 6), np.float32)
        m = C.c_size_t()
        rc = self.L.srx_mlp_samples(self.h, _fptr(x), _fptr(t), n.value, C.byref(m))
        if rc:
            raise SrxError(rc, "srx_mlp_samples")
        return x[:m.value], t[:m.value]

    def close(self) -> None:
        if self.h:
            self.L.srx_close(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hbm_stream(nbytes: int = 1 << 30, reps: int = 5) -> dict:
    """srx_hbm_stream: read + write GB/s of the best plain streaming-copy launch shape on this GPU."""
    best, mean = C.c_double(), C.c_double()
    u, b = C.c_int(), C.c_int()
    rc = lib().srx_hbm_stream(nbytes, reps, C.byref(best), C.byref(mean), C.byref(u), C.byref(b))
    if rc:
        raise SrxError(rc, "srx_hbm_stream")
    return {"best_gbs": best.value, "mean_gbs": mean.value, "unroll": u.value, "blocks_per_cu": b.value,
            "bytes_each_way": nbytes}


def fft_forward(x: np.ndarray, precision: int = PREC_FP64) -> np.ndarray:
    """Half-spectrum forward FFT of real fields x[c, n0, n1, n2] on the GPU.
    Returns complex [c, n0, n1, n2//2+1] (numpy.fft.rfftn convention)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    nc, n0, n1, n2 = x.shape
    out = np.zeros(nc * n0 * n1 * (n2 // 2 + 1) * 2)
    rc = lib().srx_fft_forward(_dptr(x), nc, n0, n1, n2, precision, _dptr(out))
    if rc:
        raise SrxError(rc, "srx_fft_forward")
    return out.view(np.complex128).reshape(nc, n0, n1, n2 // 2 + 1)


def fft_inverse(X: np.ndarray, n2: int, precision: int = PREC_FP64, scale: float = 1.0) -> np.ndarray:
    """Real inverse of fft_forward: returns scale * irfftn-without-normalisation."""
    X = np.ascontiguousarray(X, dtype=np.complex128)
    nc, n0, n1, h = X.shape
    xin = X.view(np.float64).ravel().copy()
    out = np.zeros(nc * n0 * n1 * n2)
    rc = lib().srx_fft_inverse(_dptr(xin), nc, n0, n1, n2, precision, scale, _dptr(out))
    if rc:
        raise SrxError(rc, "srx_fft_inverse")
    return out.reshape(nc, n0, n1, n2)


def mlp_train(x: np.ndarray, t: np.ndarray, blob: np.ndarray | None = None, **kw):
    """Train the learned Newton start on the GPU.  Returns (blob, loss per epoch).  kw: fields of
    srx_mlp_train_opts (epochs, batch, lr, lr_final, seed); with `blob` given, training continues
    from it (init 0)."""
    L = lib()
    o = MlpTrainOpts()
    L.srx_mlp_default_train_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    assert x.ndim == 2 and x.shape[1] == 24 and t.shape == (x.shape[0], 6)
    if blob is None:
        b = np.zeros(MLP_BLOB, np.float32)
        o.init = 1
    else:
        b = np.array(blob, dtype=np.float32).ravel()
        o.init = 0
    hist = np.zeros(max(o.epochs, 1), np.float32)
    rc = L.srx_mlp_train(_fptr(x), _fptr(t), x.shape[0], C.byref(o), _fptr(b), _fptr(hist))
    if rc:
        raise SrxError(rc, "srx_mlp_train")
    return b, hist


def mlp_grad(blob: np.ndarray, x: np.ndarray, t: np.ndarray):
    """(gradient [6150], loss) of the normalised mean squared error, one GPU pass."""
    b = np.ascontiguousarray(blob, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    g = np.zeros(MLP_NPARAM, np.float32)
    loss = C.c_double()
    rc = lib().srx_mlp_grad(_fptr(b), _fptr(x), _fptr(t), x.shape[0], _fptr(g), C.byref(loss))
    if rc:
        raise SrxError(rc, "srx_mlp_grad")
    return g, loss.value


def mlp_forward(blob: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Network outputs [n][6] for raw features [n][24] (matrix-core forward pass)."""
    b = np.ascontiguousarray(blob, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros((x.shape[0], 6), np.float32)
    rc = lib().srx_mlp_forward(_fptr(b), _fptr(x), x.shape[0], _fptr(y))
    if rc:
        raise SrxError(rc, "srx_mlp_forward")
    return y


def run_ensemble(decks, workdirs=None, concurrency: int = 4, **opt) -> list[dict]:
    """srx_run_ensemble: every increment of each deck on this GPU, `concurrency` RVEs at a time on
    their own streams, or (concurrency 0) all in lockstep with the per-voxel kernels of all members
    in one launch each.
    Returns one dict per deck (rc, increments, voxels, setup_seconds, run_seconds, run_phase_seconds
    (wall time of all members' increments, setup excluded), last report);
    raises SrxError if any member failed."""
    n = len(decks)
    arr = (C.c_char_p * n)(*[d.encode() for d in decks])
    wds = (C.c_char_p * n)(*[w.encode() for w in workdirs]) if workdirs else None
    res = (EnsembleResult * n)()
    o = options(**opt)
    rc = lib().srx_run_ensemble(arr, wds, n, C.byref(o), concurrency, res)
    out = [{"index": r.index, "rc": r.rc, "increments": r.increments, "voxels": r.voxels,
            "setup_seconds": r.setup_seconds, "run_seconds": r.run_seconds,
            "run_phase_seconds": r.run_phase_seconds, "last": r.last.as_dict()} for r in res]
    if rc:
        raise SrxError(rc, "srx_run_ensemble")
    return out
