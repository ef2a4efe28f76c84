"""srx_fft_forward_noalloc (the allocation-free variant of srx_fft_forward) through the C ABI:
device input and output buffers from the caller, spectra and twiddles in the library's static
device array.  Two calls in a row, the second smaller and in the other precision, each against
srx_fft_forward on the same input (1e-12 of the spectrum's largest magnitude in FP64, 1e-5 in mixed
precision, where both compute in FP32 and may differ in the twiddles' last bits), plus the
rejection of a problem larger than the static array."""
import ctypes as C

import numpy as np
import pytest

import srx_amd

pytestmark = pytest.mark.gpu

HIP_H2D, HIP_D2H = 1, 2
SCRATCH_BYTES = 32 << 20   # SRX_NOALLOC_SCRATCH_BYTES


def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DevBuf:
    def __init__(self, h, nbytes):
        self.h = h
        self.p = C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), nbytes) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert self.h.hipMemcpy(self.p, a.ctypes.data, a.nbytes, HIP_H2D) == 0

    def get(self, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, out.nbytes, HIP_D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)


def noalloc_forward(h, x, n, prec):
    L = srx_amd.lib()
    L.srx_fft_forward_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
    ncomp = x.shape[0]
    H = n[2] // 2 + 1
    din = DevBuf(h, x.nbytes)
    dout = DevBuf(h, ncomp * n[0] * n[1] * H * 16)
    din.put(x)
    rc = L.srx_fft_forward_noalloc(din.p, ncomp, n[0], n[1], n[2], prec, dout.p)
    out = dout.get((ncomp, n[0], n[1], H, 2)) if rc == 0 else None
    din.free()
    dout.free()
    return rc, out


def host_forward(x, n, prec):
    return srx_amd.fft_forward(x, precision=prec)


@pytest.mark.parametrize("order", ["large_then_small"])
def test_noalloc_forward_matches_srx_fft_forward_twice(order):
    h = hip()
    rng = np.random.default_rng(7)
    cases = [((24, 20, 32), 6, srx_amd.PREC_FP64, 1e-12), ((8, 12, 10), 3, srx_amd.PREC_MIXED, 1e-5)]
    for n, ncomp, prec, tol in cases:
        x = rng.standard_normal((ncomp,) + n)
        print("noalloc: n=%s ncomp=%d precision=%d" % (n, ncomp, prec), flush=True)
        rc, got = noalloc_forward(h, x, n, prec)
        assert rc == 0, rc
        want = host_forward(x, n, prec)
        g = got[..., 0] + 1j * got[..., 1]
        w = want.reshape(g.shape)
        assert np.abs(g - w).max() <= tol * np.abs(w).max()
        # and the transform itself (numpy), at the precision's accuracy
        ref = np.fft.rfftn(x, axes=(1, 2, 3))
        assert np.abs(g - ref).max() <= (1e-12 if prec == srx_amd.PREC_FP64 else 1e-5) * np.abs(ref).max()


def test_noalloc_rejects_a_problem_beyond_its_scratch():
    h = hip()
    # 9 FP64 components of 96^3: two half spectra of 9*49*96*96*16 B > 32 MiB
    n = (96, 96, 96)
    x = np.zeros((1,) + n)
    L = srx_amd.lib()
    L.srx_fft_forward_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
    din = DevBuf(h, 9 * x.nbytes)
    dout = DevBuf(h, 16)
    rc = L.srx_fft_forward_noalloc(din.p, 9, n[0], n[1], n[2], srx_amd.PREC_FP64, dout.p)
    din.free()
    dout.free()
    assert rc == 2011   # SRX_E_SCRATCH
    assert L.srx_strerror(2011).decode().startswith("problem larger than")
