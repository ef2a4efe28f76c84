"""GPU solver parity against the CPU oracle through the C ABI (srx_open/srx_step/srx_get_field).

Tolerances: the GPU and the oracle evaluate the same algorithm in a different floating-point
order (crystal-frame Newton with an LDL^T solve vs the reference's sample-frame LU inverse;
radix-4 vs radix-2 FFTs), so per-voxel fields agree to ~1e-12 relative after a few increments.
Tests with a fixed number of outer iterations compare at 1e-9 relative; converged runs use the
deck's Err and compare at a few times Err (the iteration count may then differ by one).
"""
import os
import shutil

import numpy as np
import pytest

import srx_amd
from srx_amd import inputs

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_inputs")


def make(tmp_path, n=(16, 16, 16), grains=8, seed=1, nsteps=4, **kw):
    return inputs.make_case(str(tmp_path), n=n, grains=grains, seed=seed, nsteps=nsteps, timestep=0.1, **kw)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def compare_fields(g, o, tol, names=("sig", "disgrad", "eps", "edot", "rho_s", "rho_m", "gamdot", "tranmat",
                                      "gam_acum", "rho_forest", "rho_inplane", "rho_dipole", "velgrad")):
    for f in names:
        e = rel(g.field(f), o.field(f))
        assert e < tol, (f, e)


def compare_reports(rg, ro, tol):
    for k in ("s_vm", "s_vm1", "e_vm", "evmp", "d_vm", "dvmp", "rho_s_avg", "rho_m_avg", "rho_forest_avg",
              "rho_inplane_avg", "rho_dipole_avg"):
        assert abs(rg[k] - ro[k]) <= tol * max(abs(ro[k]), 1e-300), (k, rg[k], ro[k])
    np.testing.assert_allclose(rg["sig_avg"], ro["sig_avg"], rtol=0, atol=tol * np.abs(ro["sig_avg"]).max())


def test_fixed_iterations_match_oracle(tmp_path, oracle_lib):
    deck = make(tmp_path, nsteps=4)
    o = oracle_lib.Oracle(deck, fixed_iters=6)
    g = srx_amd.Solver(deck, fixed_iters=6)
    for step in range(4):
        ro, rg = o.step(), g.step()
        assert rg["iters"] == ro["iters"] == 6
        assert rg["nr_iters_max"] == ro["nr_iters_max"]
        assert rg["nr_iters_total"] == ro["nr_iters_total"]
        assert rg["nr_wave_iters"] == ro["nr_wave_iters"] > 0
        compare_reports(rg, ro, 1e-10)
        compare_fields(g, o, 1e-9)


@pytest.mark.parametrize("precision", [srx_amd.PREC_FP64, srx_amd.PREC_MIXED])
def test_odd_voxel_count_matches_oracle(tmp_path, oracle_lib, precision):
    """A grid of odd extents (15 x 9 x 5 = 675 voxels: an odd count, so the Newton kernel stages its
    inputs with 4-byte copies, and a last 64-voxel block of 35) at fixed outer iterations, through
    increments where the texture rotates: FP64 to 1e-9 with the oracle's exact Newton iteration
    counts, mixed precision to the mixed bounds (1e-4 on the macroscopic stress, 1e-5 on rho_s)."""
    deck = make(tmp_path, n=(15, 9, 5), grains=6, seed=21, nsteps=4)
    o = oracle_lib.Oracle(deck, fixed_iters=4)
    g = srx_amd.Solver(deck, fixed_iters=4, precision=precision)
    for step in range(4):
        ro, rg = o.step(), g.step()
        assert rg["iters"] == ro["iters"] == 4
        if precision == srx_amd.PREC_FP64:
            assert rg["nr_iters_total"] == ro["nr_iters_total"]
            compare_reports(rg, ro, 1e-10)
            compare_fields(g, o, 1e-9)
        else:
            so, sm = np.array(ro["sig_avg"]), np.array(rg["sig_avg"])
            assert np.linalg.norm(sm - so) / np.linalg.norm(so) < 1e-4
            assert rel(g.field("rho_s"), o.field("rho_s")) < 1e-5
    g.close()
    o.close()


def test_fewer_slip_systems_match_oracle(tmp_path, oracle_lib):
    """A slip mode of 7 of the 12 systems (an odd count: the Newton kernel's last pair of systems
    holds one padding system, and the end-of-increment kernel's system loop and its one-ahead
    loads stop at 7) at fixed outer iterations, through increments where the texture rotates."""
    deck = make(tmp_path, n=(16, 12, 10), grains=6, seed=22, nsteps=4, nsys=7)
    o = oracle_lib.Oracle(deck, fixed_iters=4)
    g = srx_amd.Solver(deck, fixed_iters=4)
    for step in range(4):
        ro, rg = o.step(), g.step()
        assert rg["nr_iters_total"] == ro["nr_iters_total"]
        compare_reports(rg, ro, 1e-10)
        compare_fields(g, o, 1e-9)
    g.close()
    o.close()


def test_converged_increments_match_oracle(tmp_path, oracle_lib):
    """Increments iterated to the deck's Err = 2e-5, including istep >= 2 where the texture
    (and with it the crystal frame of the Newton solve) evolves."""
    deck = make(tmp_path, n=(16, 16, 16), grains=8, seed=4, nsteps=5)
    o = oracle_lib.Oracle(deck)
    g = srx_amd.Solver(deck)
    for step in range(5):
        ro, rg = o.step(), g.step()
        assert rg["converged"] and ro["converged"]
        assert abs(rg["iters"] - ro["iters"]) <= 1, (rg["iters"], ro["iters"])
        compare_reports(rg, ro, 1e-4)
    if rg["iters"] == ro["iters"]:
        compare_fields(g, o, 1e-6)


def test_drop_in_check_64cubed_125_grains(tmp_path, oracle_lib):
    """The drop-in correctness configuration: 64^3 voxels, 125 Voronoi grains, random texture,
    three increments (the texture rotates from the third) with fixed outer iterations, every field
    and report value against the oracle on identical inputs."""
    deck = make(tmp_path, n=(64, 64, 64), grains=125, seed=11, nsteps=3)
    o = oracle_lib.Oracle(deck, fixed_iters=6)
    g = srx_amd.Solver(deck, fixed_iters=6)
    for step in range(3):
        ro, rg = o.step(), g.step()
        assert rg["nr_iters_total"] == ro["nr_iters_total"]
        assert rg["nr_iters_max"] == ro["nr_iters_max"]
        compare_reports(rg, ro, 1e-10)
        compare_fields(g, o, 1e-9)
    g.close()


def test_twenty_converged_increments_16cubed(tmp_path, oracle_lib):
    """The CPU-runnable configuration: 16^3, 8 grains, 20 increments to the deck's Err; every
    increment's report within a few times Err of the oracle's, at most one outer iteration apart."""
    deck = make(tmp_path, n=(16, 16, 16), grains=8, seed=1, nsteps=20)
    o = oracle_lib.Oracle(deck)
    with srx_amd.Solver(deck) as g:
        for step in range(20):
            ro, rg = o.step(), g.step()
            assert rg["converged"] and ro["converged"]
            assert abs(rg["iters"] - ro["iters"]) <= 1, (step, rg["iters"], ro["iters"])
            compare_reports(rg, ro, 1e-4)


def test_reference_deck_two_crystalline_phases_2d(tmp_path, oracle_lib):
    """The reference's own deck: 256x256x1 (a one-voxel z extent), both phases crystalline with
    different slip-mode files, all six velocity components prescribed."""
    for f in ("input", "Cu_p.sx", "Cu_p_hard.sx"):
        shutil.copy(os.path.join(GOLD, f), tmp_path / f)
    n = (256, 256, 1)
    eul, gid = inputs.voronoi(n, 12, seed=9)
    phase = np.where(np.arange(n[0] * n[1]) // n[1] < 128, 1, 2)
    inputs.write_mesh(str(tmp_path / "Parent_Twin_Hori.txt"), n, eul, gid, phase)
    deck = str(tmp_path / "input")
    o = oracle_lib.Oracle(deck, fixed_iters=3)
    g = srx_amd.Solver(deck, fixed_iters=3)
    for step in range(3):
        ro, rg = o.step(), g.step()
        compare_reports(rg, ro, 1e-10)
        compare_fields(g, o, 1e-9, names=("sig", "disgrad", "eps", "rho_s", "rho_m", "gamdot", "tranmat"))


def test_bicrystal_out888(tmp_path, oracle_lib):
    """The 8x8x8 bicrystal of the fftw3 variant (CRLF mesh, y tension, stress-controlled 11 and
    33, tau0xf = 0), run with the top-level reference's algorithm."""
    for f in ("input", "Cu_p.sx", "out888.txt"):
        shutil.copy(os.path.join(GOLD, "fftw3", f), tmp_path / f)
    deck = str(tmp_path / "input")
    o = oracle_lib.Oracle(deck, fixed_iters=8)
    g = srx_amd.Solver(deck, fixed_iters=8)
    for step in range(4):
        ro, rg = o.step(), g.step()
        compare_reports(rg, ro, 1e-10)
        compare_fields(g, o, 1e-9)


def test_elastbc0_dc_term(tmp_path, oracle_lib):
    """ElastBC 0: the k = 0 Green operator is -S0 (init.c:558-569)."""
    deck = make(tmp_path, n=(8, 8, 8), grains=3, seed=2, nsteps=3, elast_bc=0)
    o = oracle_lib.Oracle(deck, fixed_iters=5)
    g = srx_amd.Solver(deck, fixed_iters=5)
    for step in range(3):
        ro, rg = o.step(), g.step()
        compare_reports(rg, ro, 1e-10)
        compare_fields(g, o, 1e-9, names=("sig", "disgrad", "eps", "rho_s"))


def test_mixed_precision_error_vs_fp64(tmp_path):
    """FP32 FFT/Green and FP32-stored Newton constants inside the FP64 fixed point: macroscopic
    stress within 1e-4 relative of the FP64 GPU path, per-voxel von Mises stress within 1e-3 of its
    maximum, slip-system rho_s within 1e-5 relative (the stated tolerances; this case measures
    about 1.3e-6 for rho_s)."""
    deck = make(tmp_path, n=(32, 32, 32), grains=20, seed=6, nsteps=4)
    a = srx_amd.Solver(deck)
    b = srx_amd.Solver(deck, precision=srx_amd.PREC_MIXED)
    for step in range(4):
        ra, rb = a.step(), b.step()
        sa, sb = np.array(ra["sig_avg"]), np.array(rb["sig_avg"])
        assert np.linalg.norm(sb - sa) <= 1e-4 * np.linalg.norm(sa)
    def vm(s):
        s = s.reshape(-1, 3, 3)
        d = s - np.trace(s, axis1=1, axis2=2)[:, None, None] * np.eye(3) / 3
        return np.sqrt(1.5 * (d * d).sum(axis=(1, 2)))
    va, vb = vm(a.field("sig")), vm(b.field("sig"))
    assert np.abs(va - vb).max() <= 1e-3 * va.max()
    assert rel(b.field("rho_s"), a.field("rho_s")) < 1e-5


def von_mises(s):
    s = s.reshape(-1, 3, 3)
    d = s - np.trace(s, axis1=1, axis2=2)[:, None, None] * np.eye(3) / 3
    return np.sqrt(1.5 * (d * d).sum(axis=(1, 2)))


@pytest.mark.parametrize("n,grains,seed", [((32, 32, 32), 20, 6), ((15, 9, 5), 6, 21)])
def test_mixed_precision_against_the_oracle(tmp_path, oracle_lib, n, grains, seed):
    """The mixed path against the FP64 oracle, the errors the mixed-precision configuration asks
    for: the stress-strain curve (macroscopic stress and e_vm of every converged increment) within
    1e-4 relative, per-voxel von Mises stress within 1e-3 of its maximum, the slip-system
    dislocation densities within 1e-5 relative.  675 voxels: the FP32 Newton constants' pitch is
    padded (676) and the last Newton block holds 35 voxels."""
    deck = make(tmp_path, n=n, grains=grains, seed=seed, nsteps=5)
    o = oracle_lib.Oracle(deck)
    with srx_amd.Solver(deck, precision=srx_amd.PREC_MIXED) as g:
        for step in range(5):
            ro, rg = o.step(), g.step()
            assert rg["converged"] and ro["converged"]
            assert abs(rg["iters"] - ro["iters"]) <= 1, (step, rg["iters"], ro["iters"])
            sa, sb = np.array(ro["sig_avg"]), np.array(rg["sig_avg"])
            assert np.linalg.norm(sb - sa) <= 1e-4 * np.linalg.norm(sa), step
            assert abs(rg["e_vm"] - ro["e_vm"]) <= 1e-4 * abs(ro["e_vm"]), step
        vo, vg = von_mises(o.field("sig")), von_mises(g.field("sig"))
        assert np.abs(vg - vo).max() <= 1e-3 * vo.max()
        assert rel(g.field("rho_s"), o.field("rho_s")) < 1e-5


@pytest.mark.parametrize("n,seed,fixed", [((8, 8, 8), 7, 6), ((16, 16, 16), 3, 6), ((8, 8, 8), 7, 0),
                                          ((12, 10, 8), 11, 0)])
def test_output_files_match_oracle(tmp_path, oracle_lib, n, seed, fixed):
    """vm.out / dd.out / err.out written by the GPU path are the oracle's files byte for byte (the
    reference's headers and %e columns), at a fixed number of outer iterations and converged."""
    deck = make(tmp_path, n=n, grains=3, seed=seed, nsteps=3)
    go, oo = tmp_path / "gpu", tmp_path / "orc"
    go.mkdir()
    oo.mkdir()
    for f in ("Cu_p.sx", "mesh.txt"):
        shutil.copy(tmp_path / f, go / f)
        shutil.copy(tmp_path / f, oo / f)
    g = srx_amd.Solver(deck, workdir=str(go), write_files=1, fixed_iters=fixed)
    o = oracle_lib.Oracle(deck, workdir=str(oo), write_files=1, fixed_iters=fixed)
    for _ in range(3):
        g.step()
        o.step()
    g.close()
    o.close()
    for name in ("vm.out", "dd.out", "err.out"):
        A = open(go / name).read()
        B = open(oo / name).read()
        assert len(A.splitlines()) == 4, name
        assert A == B, name


def test_nan_is_reported_not_aborted(tmp_path):
    """0.1% strain in one increment from the elastic predictor overflows the plain Newton solve
    (the reference asserts, constitutive.c:683); the C ABI returns SRX_E_NAN instead."""
    deck = inputs.make_case(str(tmp_path), n=(8, 8, 8), grains=4, seed=1, nsteps=2, timestep=1.0)
    g = srx_amd.Solver(deck)
    with pytest.raises(srx_amd.SrxError) as e:
        g.step()
    assert e.value.code == 3001


def test_warm_start_matches_oracle(tmp_path, oracle_lib):
    """warm_start 1 (extrapolated Newton start, no counterpart in the reference): the GPU and the
    oracle run the same algorithm."""
    deck = make(tmp_path, n=(16, 16, 16), grains=8, seed=12, nsteps=4)
    o = oracle_lib.Oracle(deck, fixed_iters=6, warm_start=1)
    g = srx_amd.Solver(deck, fixed_iters=6, warm_start=1)
    for step in range(4):
        ro, rg = o.step(), g.step()
        compare_reports(rg, ro, 1e-10)
        compare_fields(g, o, 1e-9)


def test_warm_start_saves_newton_iterations(tmp_path):
    """Converged increments with and without the extrapolated start: the same macroscopic response
    (the Newton tolerance is 1e-10) with fewer Newton iterations."""
    deck = make(tmp_path, n=(24, 24, 24), grains=12, seed=13, nsteps=5)
    cold = srx_amd.Solver(deck)
    warm = srx_amd.Solver(deck, warm_start=1)
    nr_cold = nr_warm = 0
    for step in range(5):
        rc, rw = cold.step(), warm.step()
        np.testing.assert_allclose(rw["sig_avg"], rc["sig_avg"], rtol=0, atol=1e-8 * np.abs(rc["sig_avg"]).max())
        assert rw["iters"] == rc["iters"]
        if step >= 2:
            nr_cold += rc["nr_iters_total"]
            nr_warm += rw["nr_iters_total"]
    assert nr_warm < 0.9 * nr_cold, (nr_warm, nr_cold)
    assert rel(warm.field("rho_s"), cold.field("rho_s")) < 1e-8


@pytest.mark.parametrize("n", [(64, 6, 4), (128, 4, 6), (256, 2, 4)])
def test_wavefront_green_pass_matches_oracle(tmp_path, oracle_lib, n):
    """x extents of 64, 128 and 256 run the Green x pass one wavefront per row (1, 2 and 4 points
    per lane, cross-lane radix-2 stages); FP64 parity with the oracle's full complex FFT, with an
    elastic-BC-0 deck so the k = 0 term is exercised too."""
    deck = make(tmp_path, n=n, grains=4, seed=15, nsteps=3, elast_bc=0)
    o = oracle_lib.Oracle(deck, fixed_iters=5)
    g = srx_amd.Solver(deck, fixed_iters=5, reserved=[0, 0, 1])   # default tiles: the wavefront pass
    for step in range(3):
        ro, rg = o.step(), g.step()
        assert rg["nr_iters_total"] == ro["nr_iters_total"]
        compare_reports(rg, ro, 1e-10)
        compare_fields(g, o, 1e-9)


@pytest.mark.parametrize("creep", [False, True])
def test_device_setup_matches_oracle(tmp_path, oracle_lib, creep):
    """Orientations, the initial stress C_gr : (Udot dt) (plus the creep stress) and the averaged
    stiffness, computed on the device at srx_open, against the oracle's setup (init.c:233-301,
    444-516) before any increment."""
    kw = dict(extra={"VelGrad_BC_Flag": "0 0 0 0 0 0", "VelGrad_BC": "0 0 0 0 0 0",
                     "Stress_BC_Flag": "1 1 1 1 1 1", "Stress_BC": "0. 0. 50. 0. 0. 0."}) if creep else {}
    deck = make(tmp_path, n=(12, 10, 8), grains=5, seed=16, nsteps=1, **kw)
    o = oracle_lib.Oracle(deck)
    g = srx_amd.Solver(deck)
    for name in ("tranmat", "sig"):
        a, b = g.field(name), o.field(name)
        assert np.abs(a - b).max() <= 1e-13 * max(np.abs(b).max(), 1.0), name
    g.close()
    o.close()
