"""GND option on the GPU (srx_options.gnd = 1, a build of the reference with DD_GND) against the CPU
oracle, whose GND steps tests/test_oracle_gnd.py pins.  Fixed outer iterations compare per-voxel
fields at 1e-9 relative, as the solver parity tests do; converged increments at a few times Err."""
import numpy as np
import pytest

import srx_amd
from srx_amd import inputs

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


FIELDS = ("sig", "disgrad", "eps", "rho_s", "rho_m", "rho_F", "rho_P", "gamdot", "rho_g1", "rho_g2", "rho_tot",
          "rho_forest", "rho_dipole", "tranmat")


@pytest.mark.parametrize("n,grains", [((16, 12, 10), 6), ((24, 24, 24), 12)])
def test_gnd_fixed_iterations_match_oracle(tmp_path, oracle_lib, n, grains):
    deck = inputs.make_case(str(tmp_path), n=n, grains=grains, seed=4, nsteps=6, timestep=0.1)
    o = oracle_lib.Oracle(deck, fixed_iters=8, gnd=1)
    g = srx_amd.Solver(deck, fixed_iters=8, gnd=1)
    for step in range(6):
        ro, rg = o.step(), g.step()
        assert rg["nr_iters_total"] == ro["nr_iters_total"]
        for k in ("s_vm", "rho_s_avg", "rho_m_avg", "rho_F_avg", "rho_P_avg", "rho_g_avg", "rho_dipole_avg"):
            assert abs(rg[k] - ro[k]) <= 1e-10 * max(abs(ro[k]), 1e-300), (step, k, rg[k], ro[k])
        for f in FIELDS:
            e = rel(g.field(f), o.field(f))
            assert e < 1e-9, (step, f, e)
    assert ro["rho_g_avg"] > 0.1
    g.close()


def test_gnd_converged_increments_match_oracle(tmp_path, oracle_lib):
    deck = inputs.make_case(str(tmp_path), n=(16, 16, 16), grains=8, seed=4, nsteps=5, timestep=0.1)
    o = oracle_lib.Oracle(deck, gnd=1)
    with srx_amd.Solver(deck, gnd=1, precision=srx_amd.PREC_MIXED) as g:
        for step in range(5):
            ro, rg = o.step(), g.step()
            assert rg["converged"] and ro["converged"]
            assert abs(rg["iters"] - ro["iters"]) <= 1
            for k in ("s_vm", "rho_g_avg", "rho_F_avg"):
                assert abs(rg[k] - ro[k]) <= 1e-4 * max(abs(ro[k]), 1e-300), (step, k, rg[k], ro[k])


def test_gnd_changes_the_response_and_is_off_by_default(tmp_path):
    deck = inputs.make_case(str(tmp_path), n=(16, 16, 16), grains=8, seed=4, nsteps=5, timestep=0.1)
    with srx_amd.Solver(deck, fixed_iters=8) as a, srx_amd.Solver(deck, fixed_iters=8, gnd=1) as b:
        for _ in range(5):
            ra, rb = a.step(), b.step()
        assert ra["rho_g_avg"] == 0.0 and rb["rho_g_avg"] > 0.0
        assert rb["rho_F_avg"] > ra["rho_F_avg"]
        with pytest.raises(srx_amd.SrxError):
            a.field("rho_g1")
    with pytest.raises(srx_amd.SrxError):
        srx_amd.Solver(deck, gnd=1, warm_start=1)
