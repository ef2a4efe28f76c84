"""Field dumps (.iout, io.c:322-941 at evolution.c:1456-1483) and the derived fields they use
(fluctuation, rho_tot), GPU path against the oracle's fields through the C ABI."""
import os

import numpy as np
import pytest

import srx_amd
from srx_amd import inputs, iout
from oracle.oracle import dptr, lib

pytestmark = pytest.mark.gpu

SYM = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def euler_deg(tranmat):
    """Angles the reference writes: EulerToTransMatrix opt 1 of the transposed TranMat_xt2sa."""
    out = np.zeros((len(tranmat), 3))
    for v, t in enumerate(tranmat):
        a = np.ascontiguousarray(t.reshape(3, 3).T).ravel().copy()
        x = [np.zeros(1), np.zeros(1), np.zeros(1)]
        lib().orc_euler_to_matrix(dptr(x[0]), dptr(x[1]), dptr(x[2]), dptr(a), 1)
        out[v] = [x[0][0], x[1][0], x[2][0]]
    return out


def test_iout_dumps_match_oracle_fields(tmp_path, oracle_lib):
    n = (8, 6, 10)
    deck = inputs.make_case(str(tmp_path), n=n, grains=5, seed=11, nsteps=3, timestep=0.1, print_control=(1, 2))
    g = srx_amd.Solver(deck, write_fields=1, fixed_iters=5)
    o = oracle_lib.Oracle(deck, fixed_iters=5)
    N = n[0] * n[1] * n[2]
    for istep in range(3):
        g.step()
        o.step()
        names = sorted(f for f in os.listdir(tmp_path) if f.endswith(".iout"))
        if istep % 2:
            continue
        assert "tex_afterDRX_S%06d.iout" % istep in names
        sig, eps, dg = o.field("sig"), o.field("eps"), o.field("disgrad")
        for stem, want in (("eps", eps), ("sig_afterDRX", sig)):
            got = iout.read_tensor(stem, istep, str(tmp_path))
            exp = np.stack([want[:, 3 * a + b] for a, b in SYM], axis=1)
            assert got.shape == (N, 6)
            assert np.abs(got - exp).max() <= 1e-9 * np.abs(exp).max(), stem
        els = iout.read_tensor("els", istep, str(tmp_path))
        exp = np.stack([(dg[:, 3 * a + b] + dg[:, 3 * b + a]) / 2 - eps[:, 3 * a + b] for a, b in SYM], axis=1)
        assert np.abs(els - exp).max() <= 1e-9 * np.abs(exp).max()
        rho = iout.read_scalar(str(tmp_path / ("rho_afterDRX_SSD_S%06d.iout" % istep)))
        assert rel(rho, o.field("rho_s").sum(axis=1)) < 1e-9
        tex = iout.read_tex(str(tmp_path / ("tex_afterDRX_S%06d.iout" % istep)))
        assert tex.size == N and os.path.getsize(tmp_path / ("tex_afterDRX_S%06d.iout" % istep)) == 48 * N
        np.testing.assert_allclose(tex["angle"], euler_deg(o.field("tranmat")), rtol=0, atol=1e-7)
        idx = np.arange(N)
        np.testing.assert_array_equal(tex["coord"][:, 0], idx // (n[1] * n[2]) + 1)
        np.testing.assert_array_equal(tex["coord"][:, 1], (idx // n[2]) % n[1] + 1)
        np.testing.assert_array_equal(tex["coord"][:, 2], idx % n[2] + 1)
        np.testing.assert_array_equal(tex["jgr"], o.field("grain").astype(int))
        np.testing.assert_array_equal(tex["jph"], o.field("phase").astype(int))
        assert (tex["buff"] == 0).all()
        for stem, fld in (("slip_afterDRX", "trial_gam_acum"), ("dislocation_afterDRX", "rho_s")):
            path = tmp_path / ("%s_S%04d.iout" % (stem, istep))
            raw = np.fromfile(path, dtype=np.uint8).reshape(N, 104)
            assert (raw[:, 4:8] == 0).all()
            rec = iout.read_sys(str(path))
            np.testing.assert_array_equal(rec["jgr"], o.field("grain").astype(int))
            assert rel(rec["ss"], o.field(fld)) < 1e-9, stem
        pos = iout.read_position(str(tmp_path / ("new_position_afterDRX_S%04d.iout" % istep)))
        fl = o.field("fluct")
        X = np.stack([idx // (n[1] * n[2]), (idx // n[2]) % n[1], idx % n[2]], axis=1).astype(float)
        D = dg.reshape(N, 3, 3)
        newpos = X + np.einsum("vij,vj->vi", D, X)
        exp = newpos - fl.reshape(N, 3, 3)[:, :, 2] - X
        np.testing.assert_allclose(pos, exp, rtol=0, atol=1e-9 * max(1.0, np.abs(exp).max()))
    assert not any("S000001" in f for f in os.listdir(tmp_path))


def test_fluctuation_and_rho_tot_match_oracle(tmp_path, oracle_lib):
    deck = inputs.make_case(str(tmp_path), n=(8, 8, 8), grains=4, seed=3, nsteps=3, timestep=0.1)
    g = srx_amd.Solver(deck, fixed_iters=5)
    o = oracle_lib.Oracle(deck, fixed_iters=5)
    for _ in range(3):
        g.step()
        o.step()
        assert rel(g.field("fluct"), o.field("fluct")) < 1e-8
        assert rel(g.field("rho_tot"), o.field("rho_tot")) < 1e-10


def test_creep_run_keeps_fluctuation(tmp_path, oracle_lib):
    """All six stress components prescribed (CREEP_FLAG, evolution.c:1119-1121): DisGrad is reset
    to the macroscopic value every increment while the fluctuation keeps accumulating."""
    deck = inputs.make_case(str(tmp_path), n=(8, 8, 8), grains=4, seed=5, nsteps=3, timestep=0.1,
                            extra={"VelGrad_BC_Flag": "0 0 0 0 0 0", "VelGrad_BC": "0 0 0 0 0 0",
                                   "Stress_BC_Flag": "1 1 1 1 1 1", "Stress_BC": "0. 0. 20. 0. 0. 0."})
    g = srx_amd.Solver(deck, fixed_iters=6)
    o = oracle_lib.Oracle(deck, fixed_iters=6)
    for _ in range(3):
        rg, ro = g.step(), o.step()
        np.testing.assert_allclose(rg["sig_avg"], ro["sig_avg"], rtol=0, atol=1e-9 * np.abs(ro["sig_avg"]).max())
        assert rel(g.field("disgrad"), o.field("disgrad")) < 1e-8
        assert rel(g.field("fluct"), o.field("fluct")) < 1e-8
