"""GPU parity at the benchmark workloads' own shapes, through the C ABI.

* 128^3, 1000 grains (the benchmarked RVE), FP64 and mixed precision, against the FP64 oracle on
  identical inputs: one increment with a fixed number of outer iterations.
* The 48-RVE ensemble of uneven sizes (16 x 32^3, 16 x 64^3, 12 x 96^3, 4 x 128^3; grain counts
  U{20..1000}, random or cube texture, strain rates log-U[1e-4, 1e-2]/s at 0.01% strain per
  increment) run by srx_run_ensemble: at fixed iterations in FP64 one member of each size against
  the oracle, and converged in mixed precision every member against itself run alone.
* 256^3, 8000 grains with the learned Newton start: the same macroscopic stress as the
  reference's start (a size-independent property: every start iterates each voxel to the same
  Newton tolerance).

Tolerances as in test_gpu_solver: fixed-iteration FP64 runs 1e-9 on fields and 1e-10 on report
values (the GPU and the oracle differ only in floating-point order); the mixed path against the
FP64 oracle 1e-4 on the macroscopic stress, 1e-3 on per-voxel von Mises stress (relative to its
maximum) and 1e-5 on slip-system densities.
"""
import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np
import pytest

import srx_amd
from srx_amd import inputs

pytestmark = pytest.mark.gpu

FIELDS = ("sig", "disgrad", "eps", "edot", "rho_s", "rho_m", "gamdot", "gam_acum", "rho_forest",
          "rho_inplane", "rho_dipole", "velgrad", "tranmat")
REPORT_KEYS = ("s_vm", "s_vm1", "e_vm", "evmp", "d_vm", "dvmp", "rho_s_avg", "rho_m_avg", "rho_forest_avg",
               "rho_inplane_avg", "rho_dipole_avg")


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def compare_reports(rg, ro, tol, keys=REPORT_KEYS):
    for k in keys:
        assert abs(rg[k] - ro[k]) <= tol * max(abs(ro[k]), 1e-300), (k, rg[k], ro[k])
    np.testing.assert_allclose(rg["sig_avg"], ro["sig_avg"], rtol=0, atol=tol * np.abs(ro["sig_avg"]).max())


def von_mises(sig9):
    s = sig9.reshape(-1, 3, 3)
    dev = s - np.trace(s, axis1=1, axis2=2)[:, None, None] * np.eye(3) / 3
    return np.sqrt(1.5 * (dev * dev).sum(axis=(1, 2)))


def test_128cubed_1000_grains_fp64_and_mixed_match_oracle(tmp_path, oracle_lib):
    """The benchmarked RVE (128^3, 1000 grains, random texture), one increment of 3 outer
    iterations: FP64 to 1e-9 in every field with the oracle's exact Newton iteration counts, the
    mixed path to the stated mixed-precision bounds."""
    n = (128, 128, 128)
    print("config 128^3/1000 grains: writing inputs", flush=True)
    deck = inputs.make_case(str(tmp_path), n=n, grains=1000, seed=20261202, nsteps=2, timestep=0.1)
    o = oracle_lib.Oracle(deck, fixed_iters=3)
    ro = o.step()
    ofield = {f: o.field(f) for f in FIELDS}
    o.close()
    print("config 128^3: oracle done, FP64 GPU run", flush=True)
    with srx_amd.Solver(deck, fixed_iters=3, precision=srx_amd.PREC_FP64) as g:
        rg = g.step()
        assert rg["iters"] == ro["iters"] == 3
        assert rg["nr_iters_total"] == ro["nr_iters_total"]
        assert rg["nr_iters_max"] == ro["nr_iters_max"]
        compare_reports(rg, ro, 1e-10)
        for f in FIELDS:
            e = rel(g.field(f), ofield[f])
            assert e < 1e-9, (f, e)
    print("config 128^3: mixed GPU run", flush=True)
    with srx_amd.Solver(deck, fixed_iters=3, precision=srx_amd.PREC_MIXED) as g:
        rm = g.step()
        sm = np.array(rm["sig_avg"])
        so = np.array(ro["sig_avg"])
        assert np.linalg.norm(sm - so) / np.linalg.norm(so) < 1e-4
        vm_g, vm_o = von_mises(g.field("sig")), von_mises(ofield["sig"])
        assert np.abs(vm_g - vm_o).max() / vm_o.max() < 1e-3
        assert rel(g.field("rho_s"), ofield["rho_s"]) < 1e-5


# --- the 48-RVE ensemble ------------------------------------------------------------------------

ENSEMBLE = [(16, 32), (16, 64), (12, 96), (4, 128)]   # (count, edge)


def ensemble_specs(seed=20261203):
    rng = np.random.default_rng(seed)
    specs = []
    for count, edge in ENSEMBLE:
        for _ in range(count):
            specs.append({"edge": edge, "grains": int(rng.integers(20, 1001)), "rate": float(10 ** rng.uniform(-4, -2)),
                          "texture": "random" if rng.random() < 0.5 else "cube", "seed": seed + len(specs)})
    return specs


def _write_member(args):
    root, k, sp = args
    d = os.path.join(root, "rve%02d_%d" % (k, sp["edge"]))
    n = (sp["edge"],) * 3
    return inputs.make_case(d, n=n, grains=sp["grains"], seed=sp["seed"], nsteps=2,
                            timestep=1e-4 / sp["rate"], rate=sp["rate"], texture=sp["texture"])


@pytest.fixture(scope="module")
def ensemble48(tmp_path_factory):
    root = str(tmp_path_factory.mktemp("ens48"))
    specs = ensemble_specs()
    print("ensemble: writing %d RVEs" % len(specs), flush=True)
    with ProcessPoolExecutor(max_workers=8) as ex:
        decks = list(ex.map(_write_member, [(root, k, sp) for k, sp in enumerate(specs)]))
    return specs, decks


def test_ensemble48_fixed_iterations_match_oracle(ensemble48, oracle_lib):
    """All 48 members in one srx_run_ensemble call, FP64, 2 increments of 3 outer iterations; the
    first member of each size against the oracle run alone on the same deck."""
    specs, decks = ensemble48
    assert sum(sp["edge"] ** 3 for sp in specs) == 23724032
    print("ensemble: FP64 fixed-iteration run of 48 members", flush=True)
    res = srx_amd.run_ensemble(decks, concurrency=8, precision=srx_amd.PREC_FP64, fixed_iters=3)
    assert all(r["rc"] == 0 and r["increments"] == 2 for r in res)
    checked = set()
    for k, sp in enumerate(specs):
        if sp["edge"] in checked:
            continue
        checked.add(sp["edge"])
        print("ensemble: oracle for member %d (%d^3, %d grains, %s)" % (k, sp["edge"], sp["grains"], sp["texture"]),
              flush=True)
        o = oracle_lib.Oracle(decks[k], fixed_iters=3)
        for _ in range(2):
            ro = o.step()
        o.close()
        rg = res[k]["last"]
        assert rg["nr_iters_total"] == ro["nr_iters_total"], k
        compare_reports(rg, ro, 1e-10)
    assert checked == {32, 64, 96, 128}


def test_ensemble48_converged_mixed_members_equal_runs_alone(ensemble48):
    """The ensemble configuration as benchmarked (mixed precision, every increment to Err): each
    member's last report is bit-identical to the same deck run alone."""
    specs, decks = ensemble48
    print("ensemble: mixed converged run of 48 members (streams, then lockstep)", flush=True)
    res = srx_amd.run_ensemble(decks, concurrency=8, precision=srx_amd.PREC_MIXED)
    lock = srx_amd.run_ensemble(decks, concurrency=0, precision=srx_amd.PREC_MIXED)
    for k in range(len(decks)):
        assert lock[k]["rc"] == 0 and lock[k]["increments"] == 2
        for key in ("iters", "nr_iters_total", "s_vm", "e_vm", "rho_s_avg", "sig_avg"):
            assert lock[k]["last"][key] == res[k]["last"][key], (k, key)
    for k, deck in enumerate(decks):
        with srx_amd.Solver(deck, precision=srx_amd.PREC_MIXED) as s:
            for _ in range(2):
                last = s.step()
        r = res[k]
        assert r["rc"] == 0 and r["increments"] == 2 and r["voxels"] == specs[k]["edge"] ** 3
        assert last["converged"] == 1
        for key in ("iters", "nr_iters_total", "s_vm", "e_vm", "rho_s_avg"):
            assert r["last"][key] == last[key], (k, key)
        assert r["last"]["sig_avg"] == last["sig_avg"], k


# --- 256^3 with the learned start -----------------------------------------------------------------

def test_256cubed_8000_grains_learned_start_reaches_the_plain_stress(tmp_path):
    """The learned-start configuration's RVE (256^3, 8000 grains, mixed precision): a network
    trained on the solver's own states of a 32^3 run gives, after two converged increments, the
    same outer iteration counts and macroscopic stress as the reference's Newton start."""
    tdeck = inputs.make_case(str(tmp_path / "train"), n=(32, 32, 32), grains=40, seed=5, nsteps=4, timestep=0.1)
    with srx_amd.Solver(tdeck, precision=srx_amd.PREC_MIXED, warm_start=1) as s:
        s.mlp_record(1, 1000000)
        for _ in range(3):
            s.step()
        x, t = s.mlp_samples()
    blob, hist = srx_amd.mlp_train(x, t, epochs=5)
    print("256^3: trained on %d samples, loss %.4f -> %.4f; writing inputs" % (len(x), hist[0], hist[-1]), flush=True)
    deck = inputs.make_case(str(tmp_path / "eval"), n=(256, 256, 256), grains=8000, seed=7, nsteps=2, timestep=0.1)
    reps = {}
    for name, ws in (("plain", 0), ("learned", 2)):
        print("256^3: %s start" % name, flush=True)
        with srx_amd.Solver(deck, precision=srx_amd.PREC_MIXED, warm_start=ws) as s:
            if ws == 2:
                s.mlp_set(blob)
            reps[name] = [s.step() for _ in range(2)]
    for a, b in zip(reps["plain"], reps["learned"]):
        assert a["converged"] == b["converged"] == 1
        assert a["iters"] == b["iters"]
        sa, sb = np.array(a["sig_avg"]), np.array(b["sig_avg"])
        assert np.abs(sa - sb).max() <= 1e-8 * np.abs(sa).max()
