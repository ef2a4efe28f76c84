"""FFT radix-plan candidates (srx_fft_plan_candidates): host logic, no GPU.  The candidate set is
checked against a brute-force enumeration of the radix multisets the tuner may time, and the
default plan against fixed known values."""
import ctypes as C
import itertools
import math

import pytest

import srx_amd

# every extent the FFT passes take: 2^a 3^b 5^c <= 1024
EXTENTS = sorted(2 ** a * 3 ** b * 5 ** c for a in range(11) for b in range(7) for c in range(5)
                 if 2 ** a * 3 ** b * 5 ** c <= 1024)


def multisets(n):
    """Radix multisets of n: 8s and 4s for the power of two with at most one 2, then the 3s and 5s
    (brute force over the counts)."""
    out = set()
    for n8, n4, n2, n3, n5 in itertools.product(range(4), range(6), range(2), range(7), range(5)):
        rad = (8,) * n8 + (4,) * n4 + (2,) * n2 + (3,) * n3 + (5,) * n5
        if math.prod(rad) == n and len(rad) <= srx_amd.FFT_MAX_STAGES:
            out.add(tuple(sorted(rad)))
    return out


@pytest.mark.parametrize("n,default", [(128, [8, 4, 4]), (256, [8, 8, 4]), (64, [8, 8]), (32, [8, 4]),
                                       (96, [8, 4, 3]), (100, [4, 5, 5]), (1024, [8, 8, 4, 4]), (2, [2]),
                                       (30, [2, 3, 5]), (16, [4, 4])])
def test_default_plan_comes_first(n, default):
    assert srx_amd.fft_plan_candidates(n)[0] == default


@pytest.mark.parametrize("n", EXTENTS)
def test_candidates_are_the_allowed_factorisations(n):
    cands = srx_amd.fft_plan_candidates(n)
    for r in cands:
        assert math.prod(r) == n
        assert set(r) <= {2, 3, 4, 5, 8}
        assert r.count(2) <= 1
    assert len({tuple(r) for r in cands}) == len(cands)             # no duplicates
    assert {tuple(sorted(r)) for r in cands} == multisets(n)       # every multiset, nothing else
    for r in cands:                                                  # both stage orders of each
        assert list(reversed(r)) in cands


@pytest.mark.parametrize("n", [0, 7, 11, 1025, 2048, -4])
def test_unsupported_extent_is_rejected(n):
    with pytest.raises(srx_amd.SrxError) as e:
        srx_amd.fft_plan_candidates(n)
    assert e.value.code == srx_amd.SRX_E_GRID


def test_small_buffer_reports_the_count():
    L = srx_amd.lib()
    cnt = C.c_int()
    buf = (C.c_int * (2 * srx_amd.FFT_MAX_STAGES))()
    assert L.srx_fft_plan_candidates(128, buf, 2, C.byref(cnt)) == srx_amd.SRX_E_SIZE
    assert cnt.value == 6
    assert list(buf[:3]) == [8, 4, 4] and buf[3] == 0
    assert L.srx_fft_plan_candidates(128, None, 0, None) == 2010   # SRX_E_ARG
