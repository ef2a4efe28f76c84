"""FFT radix plans other than the default: set by srx_set_fft_plan or chosen by the opt-in tuner
(options.reserved[3] = 1).  Another radix order only reorders the FFT's rounding, so the solver's
fields agree with the default plans to the FFT's rounding error (FP64 1e-9 of each field's scale,
mixed 1e-4; the default plans are themselves checked against the oracle in test_gpu_solver.py)."""
import numpy as np
import pytest

import srx_amd
from srx_amd import inputs

pytestmark = pytest.mark.gpu

TOL = {srx_amd.PREC_FP64: 1e-9, srx_amd.PREC_MIXED: 1e-4}


def close_fields(a, b, tol):
    for f in ("sig", "disgrad", "rho_s", "eps"):
        x, y = a.field(f), b.field(f)
        scale = max(np.abs(x).max(), 1e-300)
        assert np.abs(x - y).max() <= tol * scale, f


# (grid, plan per axis): 32 x 24 x 20 takes plans on all three axes (the LDS Green pass at n0 = 32,
# radix 3 and 5 stages); 64 x 64 x 32 runs the y pass's runtime stages instead of its compile-time
# power-of-two stages (x runs the wavefront Green pass, which takes no plan)
CASES = [((32, 24, 20), {0: [2, 4, 4], 1: [3, 2, 4], 2: [5, 4]}),
         ((64, 64, 32), {1: [2, 4, 8], 2: [2, 4, 4]})]


@pytest.mark.parametrize("precision", [srx_amd.PREC_FP64, srx_amd.PREC_MIXED])
@pytest.mark.parametrize("n,plans", CASES)
def test_set_plans_match_default_plans(tmp_path, precision, n, plans):
    print("case", n, plans, "precision", precision, flush=True)
    deck = inputs.make_case(str(tmp_path), n=n, grains=6, seed=41, nsteps=2, timestep=0.1)
    a = srx_amd.Solver(deck, precision=precision, fixed_iters=4)
    b = srx_amd.Solver(deck, precision=precision, fixed_iters=4)
    for ax, r in plans.items():
        assert r in srx_amd.fft_plan_candidates(n[ax])
        assert r != a.fft_plan(ax)
        b.set_fft_plan(ax, r)
        assert b.fft_plan(ax) == r
    passes = {0: [2], 1: [1, 3], 2: [0, 4]}   # Green x; forward and inverse y; forward and inverse z
    for ax in plans:
        assert all(b.fft_tiles()[0][p] == 0 for p in passes[ax])   # back to the default tiles
    for _ in range(2):
        ra, rb = a.step(), b.step()
        assert ra["iters"] == rb["iters"] == 4
        assert abs(ra["s_vm"] - rb["s_vm"]) <= TOL[precision] * abs(ra["s_vm"])
    close_fields(a, b, TOL[precision])
    a.close()
    b.close()


def test_set_plan_rejects_bad_plans(tmp_path):
    deck = inputs.make_case(str(tmp_path), n=(32, 24, 20), grains=4, seed=42, nsteps=1, timestep=0.1)
    s = srx_amd.Solver(deck)
    before = [s.fft_plan(ax) for ax in range(3)]
    for ax, r in [(0, [8, 8]), (1, [6, 4]), (2, [4, 4, 2]), (2, []), (3, [5, 4]), (-1, [5, 4])]:
        with pytest.raises(srx_amd.SrxError) as e:
            s.set_fft_plan(ax, r)
        assert e.value.code == 2010   # SRX_E_ARG
    assert [s.fft_plan(ax) for ax in range(3)] == before
    s.close()


def test_plan_set_between_increments_drops_the_captured_batches(tmp_path):
    """Batches captured as graphs hold the plan as a kernel argument: after srx_set_fft_plan the
    graph-replayed run must stay bit-identical to the same run enqueued launch by launch."""
    deck = inputs.make_case(str(tmp_path), n=(32, 24, 20), grains=6, seed=43, nsteps=4, timestep=0.1)
    a = srx_amd.Solver(deck)
    b = srx_amd.Solver(deck, reserved=[0, 1])
    for k in range(4):
        if k == 2:
            for s in (a, b):
                s.set_fft_plan(1, [3, 2, 4])
                s.set_fft_plan(2, [5, 4])
        ra, rb = a.step(), b.step()
        assert ra["iters"] > 8   # more than one batch per increment
        assert ra["sig_avg"] == rb["sig_avg"] and ra["nr_iters_total"] == rb["nr_iters_total"]
    np.testing.assert_array_equal(a.field("disgrad"), b.field("disgrad"))
    a.close()
    b.close()


@pytest.mark.parametrize("precision", [srx_amd.PREC_FP64, srx_amd.PREC_MIXED])
def test_tuned_plans_are_candidates_shared_and_accurate(tmp_path, precision):
    deck = inputs.make_case(str(tmp_path), n=(32, 24, 20), grains=6, seed=44, nsteps=2, timestep=0.1)
    a = srx_amd.Solver(deck, precision=precision, fixed_iters=4)
    t = srx_amd.Solver(deck, precision=precision, fixed_iters=4, reserved=[0, 0, 0, 1])
    u = srx_amd.Solver(deck, precision=precision, fixed_iters=4, reserved=[0, 0, 0, 1])
    for ax, n in enumerate((32, 24, 20)):
        assert t.fft_plan(ax) in srx_amd.fft_plan_candidates(n)
        assert u.fft_plan(ax) == t.fft_plan(ax)   # one choice per grid shape in the process
    print("tuned plans", [t.fft_plan(ax) for ax in range(3)], flush=True)
    for _ in range(2):
        a.step()
        t.step()
    close_fields(a, t, TOL[precision])
    for s in (a, t, u):
        s.close()
