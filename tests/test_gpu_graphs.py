"""Outer-iteration batches replayed from HIP graphs (the default) give bit-identical results to the
same launches enqueued one by one (options.reserved[1] = 1), over several increments, so every
captured batch kind (first batch of an increment, later batches, before and after the texture
starts to rotate) is replayed several times with host work in between."""
import numpy as np
import pytest

import srx_amd
from srx_amd import inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [srx_amd.PREC_FP64, srx_amd.PREC_MIXED])
def test_graph_replay_is_bit_identical(tmp_path, precision):
    deck = inputs.make_case(str(tmp_path), n=(16, 12, 10), grains=6, seed=31, nsteps=5, timestep=0.1)
    a = srx_amd.Solver(deck, precision=precision)
    b = srx_amd.Solver(deck, precision=precision, reserved=[0, 1])
    for _ in range(5):
        ra, rb = a.step(), b.step()
        assert ra["iters"] > 8   # more than one batch per increment
        for k in ("iters", "nr_iters_total", "nr_iters_max", "s_vm", "e_vm", "err_s", "err_e", "rho_s_avg"):
            assert ra[k] == rb[k], k
        assert ra["sig_avg"] == rb["sig_avg"]
        # host work between replays: read fields back
        for f in ("sig", "disgrad", "rho_s"):
            np.testing.assert_array_equal(a.field(f), b.field(f))
    a.close()
    b.close()


def test_graphs_rebuilt_after_new_weights(tmp_path):
    """Installing learned-start weights drops the captured batches (they hold kernel arguments)."""
    from oracle import mlp_oracle as M
    deck = inputs.make_case(str(tmp_path), n=(12, 12, 12), grains=4, seed=32, nsteps=4, timestep=0.1)
    zero = M.random_blob(2)
    zero[M.O_W3:M.NPARAM] = 0.0
    zero[M.O_TM:M.O_TS] = 0.0
    a = srx_amd.Solver(deck, warm_start=2)
    b = srx_amd.Solver(deck, warm_start=2, reserved=[0, 1])
    for s in (a, b):
        s.mlp_set(zero)
    for k in range(4):
        if k == 2:
            other = M.random_blob(3)
            for s in (a, b):
                s.mlp_set(other)
        ra, rb = a.step(), b.step()
        assert ra["nr_iters_total"] == rb["nr_iters_total"]
        assert ra["sig_avg"] == rb["sig_avg"]
    a.close()
    b.close()


@pytest.mark.parametrize("precision", [srx_amd.PREC_FP64, srx_amd.PREC_MIXED])
def test_tuned_fft_tiles_are_valid_and_bit_identical(tmp_path, precision):
    """The tile autotuner only changes how many pencils a block transforms: each pencil's
    arithmetic is the same, so results are bit-identical to the default tiles."""
    deck = inputs.make_case(str(tmp_path), n=(32, 24, 20), grains=6, seed=33, nsteps=2, timestep=0.1)
    a = srx_amd.Solver(deck, precision=precision)
    b = srx_amd.Solver(deck, precision=precision, reserved=[0, 0, 1])
    tiles, us = a.fft_tiles()
    assert all(t >= 1 for t in tiles) and all(u > 0 for u in us)
    assert b.fft_tiles()[0] == [0, 0, 0, 0, 0]
    for _ in range(2):
        ra, rb = a.step(), b.step()
        assert ra["sig_avg"] == rb["sig_avg"] and ra["nr_iters_total"] == rb["nr_iters_total"]
    np.testing.assert_array_equal(a.field("disgrad"), b.field("disgrad"))
    a.close()
    b.close()
