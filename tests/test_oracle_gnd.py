"""The GND option of the CPU oracle (a build of the reference with DD_GND, which evpFlags.h pairs
with DD_GND_FIRST_ORDER), checked against an independent numpy restatement of each step:

* Gradient_ExchangeGrainID + Gradient_ShearRate (constitutive.c:731-1115): per system, the one-sided
  or central difference of gamdot inside the grain along x, y, z, periodic, with the z- same-grain
  flag taken from the y- neighbour as the reference computes it (constitutive.c:823, 842, 860);
  zero before istep 2 (evolution.c:1346-1353, init.c:1083);
* the first-order GND update of DislocationEvolution (constitutive.c:1428-1519):
  d rho_g1 = (grad . bp) / b / L0 dt, d rho_g2 = -(grad . bb) / b / L0 dt, bp = bb x bn;
* the GND terms of rho_F and rho_P (constitutive.c:565-576, tables of init.c:1118-1141), of
  rho_tot (evolution.c:1040-1046) and of AvgDislocation's GND mean (constitutive.c:1745-1757).

The numpy side works from arrays (np.roll over the grid, vectorised sums), the oracle from
per-voxel loops; a transposed axis, a wrong neighbour, a dropped |.|, a sign or a swapped table
fails these checks.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from srx_amd import inputs

A0, L0, DT = 0.362, 0.0045, 0.1


def slip_vectors(sx_path):
    lines = open(sx_path).read().splitlines()
    k = next(i for i, l in enumerate(lines) if "Slip (n-b)" in l)
    nb = np.array([[float(t) for t in l.split()] for l in lines[k + 1:k + 13]])
    n = nb[:, :3] / np.linalg.norm(nb[:, :3], axis=1, keepdims=True)
    b = nb[:, 3:] / np.linalg.norm(nb[:, 3:], axis=1, keepdims=True)
    return n, b


def grad_gamdot(gd, grain, shape):
    """[N][12][3] gradient of gamdot[N][12] with the reference's same-grain flags."""
    g = grain.reshape(shape)
    G = gd.reshape(shape + (12,))
    out = np.zeros(shape + (12, 3))
    # (axis, axis of the "minus" flag): the z- flag compares with the y- neighbour
    for d, (ax, flag_ax) in enumerate(((0, 0), (1, 1), (2, 1))):
        c0 = (g == np.roll(g, -1, axis=ax))[..., None]          # + neighbour in the same grain
        c1 = (g == np.roll(g, 1, axis=flag_ax))[..., None]      # - side flag
        vp, vm = np.roll(G, -1, axis=ax), np.roll(G, 1, axis=ax)
        out[..., d] = np.where(~c0 & c1, G - vm, np.where(c0 & ~c1, vp - G, np.where(c0 & c1, (vp - vm) / 2.0, 0.0)))
    return out.reshape(-1, 12, 3)


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    root = str(tmp_path_factory.mktemp("gnd"))
    deck = inputs.make_case(root, n=(12, 10, 8), grains=6, seed=4, nsteps=6, timestep=DT)
    o = O.Oracle(deck, gnd=1)
    names = ("gamdot", "rho_g1", "rho_g2", "rho_s", "rho_m", "rho_F", "rho_P", "rho_tot")
    hist = [{k: o.field(k).copy() for k in names}]
    reps = []
    for _ in range(5):
        reps.append(o.step())
        hist.append({k: o.field(k).copy() for k in names})
    grain = o.field("grain")
    chi, chi1 = o.setup(4, 144).reshape(12, 12), o.setup(5, 144).reshape(12, 12)
    n, b = slip_vectors(os.path.join(root, "Cu_p.sx"))
    return dict(hist=hist, reps=reps, grain=grain, chi=chi, chi1=chi1, n=n, b=b, shape=o.n)


def test_no_gnd_before_the_third_increment(run):
    for k in (1, 2):
        assert not run["hist"][k]["rho_g1"].any() and not run["hist"][k]["rho_g2"].any()
    assert np.abs(run["hist"][3]["rho_g1"]).max() > 1e-2


@pytest.mark.parametrize("k", [3, 4, 5])
def test_gnd_update_is_the_first_order_rate_of_the_gamdot_gradient(run, k):
    h0, h1 = run["hist"][k - 1], run["hist"][k]
    grad = grad_gamdot(h1["gamdot"], run["grain"], run["shape"])
    n, b = run["n"], run["b"]
    bp = np.cross(b, n)
    bb = A0 / np.sqrt(2.0)
    d1 = np.einsum("vsd,sd->vs", grad, bp) / bb / L0 * DT
    d2 = -np.einsum("vsd,sd->vs", grad, b) / bb / L0 * DT
    scale = np.abs(d1).max()
    assert scale > 0
    np.testing.assert_allclose(h1["rho_g1"] - h0["rho_g1"], d1, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(h1["rho_g2"] - h0["rho_g2"], d2, rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("k", [4, 5])
def test_gnd_terms_of_the_forest_and_parallel_densities(run, k):
    """rho_F, rho_P of increment k come from rho_s and the GND densities at its start."""
    h0, h1 = run["hist"][k - 1], run["hist"][k]
    n, b = run["n"], run["b"]
    cos_g1 = np.abs(n @ b.T)                           # |n_i . b_j|
    cos_g2 = np.abs(n @ np.cross(n, b).T)              # |n_i . (n_j x b_j)|
    sin_g1, sin_g2 = np.sqrt(np.abs(1 - cos_g1 ** 2)), np.sqrt(np.abs(1 - cos_g2 ** 2))
    g1, g2, rs = h0["rho_g1"], h0["rho_g2"], h0["rho_s"]
    rF = np.einsum("ij,vij->vi", run["chi1"], rs[:, None, :] + np.abs(g1[:, None, :] * cos_g1) + np.abs(g2[:, None, :] * cos_g2))
    rP = np.einsum("ij,vij->vi", run["chi"], rs[:, None, :] + np.abs(g1[:, None, :] * sin_g1) + np.abs(g2[:, None, :] * sin_g2))
    assert np.abs(rF - np.einsum("ij,vj->vi", run["chi1"], rs)).max() > 1e-3   # the GND part matters
    np.testing.assert_allclose(h1["rho_F"], rF, rtol=1e-13)
    np.testing.assert_allclose(h1["rho_P"], rP, rtol=1e-13)


@pytest.mark.parametrize("k", [4, 5])
def test_gnd_terms_of_rho_tot_and_the_mean(run, k):
    h0, h1 = run["hist"][k - 1], run["hist"][k]
    tot = (h0["rho_s"] + h0["rho_m"] + np.sqrt(h0["rho_g1"] ** 2 + h0["rho_g2"] ** 2)).sum(1)
    np.testing.assert_allclose(h1["rho_tot"], tot, rtol=1e-13)
    mean = np.sqrt(h1["rho_g1"] ** 2 + h1["rho_g2"] ** 2).sum() / h1["rho_g1"].shape[0]
    assert abs(run["reps"][k - 1]["rho_g_avg"] - mean) <= 1e-12 * mean


def test_option_off_keeps_the_shipped_model(tmp_path):
    deck = inputs.make_case(str(tmp_path), n=(8, 8, 8), grains=3, seed=2, nsteps=4, timestep=DT)
    a, b = O.Oracle(deck), O.Oracle(deck, gnd=0)
    for _ in range(4):
        ra, rb = a.step(), b.step()
        assert ra["rho_g_avg"] == 0.0 and ra["s_vm"] == rb["s_vm"]
    with pytest.raises(RuntimeError):
        a.field("rho_g1")
