"""The `srx` command line tool, the drop-in for the reference's `trial <deck>` (evp.c:7-31): run
from the case directory, it writes vm.out, dd.out, err.out and drx.out there with the values of the
C ABI's step reports, and returns the reference's error code on a bad deck."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import srx_amd
from srx_amd import inputs

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "srx_amd", "srx")


def rows(path):
    with open(path) as f:
        lines = f.read().splitlines()
    return lines[0], np.array([[float(x) for x in l.split()] for l in lines[1:] if l.strip()])


@pytest.mark.parametrize("flags", [[], ["--mixed"], ["--gnd"]])
def test_cli_runs_the_deck_and_writes_the_reference_files(tmp_path, flags):
    if not os.path.exists(CLI):
        pytest.fail("srx CLI not built")
    work = str(tmp_path)
    deck = inputs.make_case(work, n=(16, 16, 16), grains=8, seed=3, nsteps=4, timestep=0.1)
    r = subprocess.run([CLI, os.path.basename(deck), "--no-fields", *flags], cwd=work, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("STEP = ") == 4
    for name in ("vm.out", "dd.out", "err.out", "drx.out"):
        assert os.path.exists(os.path.join(work, name)), name
    head, vm = rows(os.path.join(work, "vm.out"))
    assert head.split() == ["TIME", "EVM", "EVMP", "DVM", "DVMP", "SVM", "SVMSVM_1"]
    assert vm.shape == (4, 7)
    # the same run through the C ABI gives the printed columns (%e: 6 significant digits)
    opt = {"precision": srx_amd.PREC_MIXED} if "--mixed" in flags else {}
    if "--gnd" in flags:
        opt["gnd"] = 1
    with srx_amd.Solver(deck, **opt) as s:
        reps = [s.step() for _ in range(4)]
    want = np.array([[q["time_tot"], q["e_vm"], q["evmp"], q["d_vm"], q["dvmp"], q["s_vm"], q["s_vm1"]] for q in reps])
    np.testing.assert_allclose(vm, want, rtol=1e-5, atol=1e-300)
    head, dd = rows(os.path.join(work, "dd.out"))
    assert dd.shape == (4, 7)
    np.testing.assert_allclose(dd[:, 2], [q["rho_s_avg"] for q in reps], rtol=1e-5)


def test_cli_returns_the_reference_error_code(tmp_path):
    if not os.path.exists(CLI):
        pytest.fail("srx CLI not built")
    work = str(tmp_path)
    deck = inputs.make_case(work, n=(8, 8, 8), grains=2, seed=1, nsteps=1, timestep=0.1)
    os.remove(os.path.join(work, "Cu_p.sx"))      # the slip-mode file the deck names
    r = subprocess.run([CLI, os.path.basename(deck)], cwd=work, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "srx:" in r.stderr and "(code" in r.stderr
    r = subprocess.run([CLI], cwd=work, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "usage" in r.stderr
