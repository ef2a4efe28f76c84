"""GPU solver parity against the CPU oracle through the C ABI (srx_open/srx_step/srx_get_field)."""
import os

import numpy as np
import pytest

import srx_amd
from srx_amd import inputs

pytestmark = pytest.mark.gpu


def make(tmp_path, n=(16, 16, 16), grains=8, seed=1, nsteps=4, **kw):
    return inputs.make_case(str(tmp_path), n=n, grains=grains, seed=seed, nsteps=nsteps, timestep=0.1, **kw)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_fixed_iterations_match_oracle(tmp_path, oracle_lib):
    deck = make(tmp_path, nsteps=4)
    o = oracle_lib.Oracle(deck, fixed_iters=6)
    g = srx_amd.Solver(deck, fixed_iters=6)
    for step in range(4):
        ro, rg = o.step(), g.step()
        print(step, ro["sig_avg"][8], rg["sig_avg"][8], ro["s_vm"], rg["s_vm"], ro["nr_iters_max"], rg["nr_iters_max"])
        assert rg["iters"] == ro["iters"] == 6
        assert abs(rg["s_vm"] - ro["s_vm"]) <= 1e-9 * abs(ro["s_vm"])
        for f in ("sig", "disgrad", "eps"):
            e = rel(g.field(f), o.field(f))
            assert e < 1e-9, (step, f, e)
        for f in ("rho_s", "rho_m", "gamdot", "tranmat"):
            e = rel(g.field(f), o.field(f))
            assert e < 1e-8, (step, f, e)
