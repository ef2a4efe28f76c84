"""srx_run_ensemble: several RVEs of uneven size run concurrently on one GPU (one stream and host
thread per member) give exactly the results of running each deck alone."""
import os

import numpy as np
import pytest

import srx_amd
from srx_amd import inputs

pytestmark = pytest.mark.gpu

CASES = [  # (grid, grains, seed, strain rate, texture)
    ((16, 16, 16), 10, 21, 1e-3, "random"),
    ((8, 12, 10), 5, 22, 3e-3, "cube"),
    ((12, 12, 12), 7, 23, 5e-4, "random"),
    ((8, 8, 8), 3, 24, 1e-3, "cube"),
    ((9, 6, 5), 3, 25, 2e-3, "random"),   # 270 voxels: padded pitches, a ragged last Newton block
]


def make_cases(root):
    decks = []
    for k, (n, g, seed, rate, tex) in enumerate(CASES):
        d = os.path.join(root, "rve%d" % k)
        # 0.01% strain per increment at every rate
        decks.append(inputs.make_case(d, n=n, grains=g, seed=seed, nsteps=3, timestep=1e-4 / rate, rate=rate,
                                      texture=tex))
    return decks


def test_ensemble_matches_members_run_alone(tmp_path):
    decks = make_cases(str(tmp_path / "ens"))
    res = srx_amd.run_ensemble(decks, concurrency=3, write_files=1)
    # every member is set up before the timed run phase, which covers all members' increments
    phase = res[0]["run_phase_seconds"]
    assert all(r["run_phase_seconds"] == phase for r in res)
    assert phase >= max(r["run_seconds"] for r in res) > 0
    assert all(r["setup_seconds"] > 0 for r in res)
    alone = make_cases(str(tmp_path / "alone"))
    for k, deck in enumerate(alone):
        s = srx_amd.Solver(deck, write_files=1)
        last = None
        for _ in range(3):
            last = s.step()
        s.close()
        r = res[k]
        assert r["rc"] == 0 and r["increments"] == 3 and r["index"] == k
        assert r["voxels"] == int(np.prod(CASES[k][0]))
        for key in ("iters", "nr_iters_total", "s_vm", "e_vm", "rho_s_avg"):
            assert r["last"][key] == last[key], (k, key)
        assert r["last"]["sig_avg"] == last["sig_avg"]
        a = open(os.path.join(os.path.dirname(decks[k]), "vm.out")).read()
        b = open(os.path.join(os.path.dirname(deck), "vm.out")).read()
        assert a == b


@pytest.mark.parametrize("concurrency,precision,gnd", [(0, srx_amd.PREC_FP64, 0), (-2, srx_amd.PREC_FP64, 0),
                                                       (0, srx_amd.PREC_MIXED, 0), (0, srx_amd.PREC_FP64, 1)])
def test_lockstep_ensemble_matches_members_run_alone(tmp_path, concurrency, precision, gnd):
    """concurrency 0: all members advance together, their increment prologues, Newton solves,
    reductions, iteration control and ends of increment in one launch each over the member table
    (-2: two such groups, concurrently); every member's results equal its deck run alone, bit for
    bit (the same per-voxel code per member, only launched together), in both precisions and with
    the GND option."""
    decks = make_cases(str(tmp_path / "ens"))
    res = srx_amd.run_ensemble(decks, concurrency=concurrency, write_files=1, precision=precision, gnd=gnd)
    alone = make_cases(str(tmp_path / "alone"))
    for k, deck in enumerate(alone):
        s = srx_amd.Solver(deck, write_files=1, precision=precision, gnd=gnd)
        for _ in range(3):
            last = s.step()
        s.close()
        r = res[k]
        assert r["rc"] == 0 and r["increments"] == 3, (k, r["rc"])
        for key in ("iters", "nr_iters_total", "s_vm", "e_vm", "rho_s_avg", "rho_g_avg"):
            assert r["last"][key] == last[key], (k, key)
        assert r["last"]["sig_avg"] == last["sig_avg"]
        a = open(os.path.join(os.path.dirname(decks[k]), "vm.out")).read()
        b = open(os.path.join(os.path.dirname(deck), "vm.out")).read()
        assert a == b


def test_lockstep_members_with_different_increment_counts(tmp_path):
    """A member with fewer increments leaves the lockstep when it is done; the others go on."""
    decks = make_cases(str(tmp_path / "ens"))[:3]
    res = srx_amd.run_ensemble(decks, concurrency=0, nsteps_override=0)
    assert [r["increments"] for r in res] == [3, 3, 3]
    d2 = str(tmp_path / "short")
    n, g, seed, rate, tex = CASES[1]
    short = inputs.make_case(d2, n=n, grains=g, seed=seed, nsteps=1, timestep=1e-4 / rate, rate=rate, texture=tex)
    res = srx_amd.run_ensemble([decks[0], short, decks[2]], concurrency=0)
    assert [r["increments"] for r in res] == [3, 1, 3]
    assert all(r["rc"] == 0 for r in res)
    assert res[0]["last"]["sig_avg"] == srx_amd.run_ensemble([decks[0]], concurrency=0)[0]["last"]["sig_avg"]


def test_ensemble_reports_a_failing_member(tmp_path):
    decks = make_cases(str(tmp_path / "ens"))[:2]
    os.remove(os.path.join(os.path.dirname(decks[1]), "mesh.txt"))
    with pytest.raises(srx_amd.SrxError) as e:
        srx_amd.run_ensemble(decks, concurrency=2)
    assert e.value.code == 2001
