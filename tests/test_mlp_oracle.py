"""Pins the CPU restatement of the learned Newton start (oracle/mlp_oracle.py): backward pass
against central finite differences, bf16 rounding against torch, quaternion against scipy,
crystal-frame stress against closed forms, and the start rule's limits."""
import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from oracle import mlp_oracle as M


def test_backward_matches_central_differences():
    rng = np.random.default_rng(3)
    blob = M.random_blob(3, xm=rng.normal(0, 0.3, 24), xs=rng.uniform(0.5, 2, 24), tm=rng.normal(0, 1, 6),
                         ts=rng.uniform(0.5, 2, 6)).astype(np.float64)
    x = rng.normal(0, 1, (40, 24))
    t = rng.normal(0, 1, (40, 6))
    loss, g = M.loss_grad(blob, x, t, rounding=False)
    idx = np.concatenate([rng.choice(np.arange(lo, hi), 12, replace=False) for lo, hi in
                          ((M.O_W1, M.O_B1), (M.O_B1, M.O_W2), (M.O_W2, M.O_B2), (M.O_B2, M.O_W3),
                           (M.O_W3, M.O_B3))] + [np.arange(M.O_B3, M.NPARAM)])
    h = 1e-6
    fd = []
    for p in idx:
        bp, bm = blob.copy(), blob.copy()
        bp[p] += h
        bm[p] -= h
        fd.append((M.loss_grad(bp, x, t, rounding=False)[0] - M.loss_grad(bm, x, t, rounding=False)[0]) / (2 * h))
    fd = np.array(fd)
    scale = np.abs(g[idx]).max()
    np.testing.assert_allclose(g[idx], fd, rtol=0, atol=1e-6 * scale)
    assert loss > 0


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-20, 100), rng.normal(0, 1e20, 100)]).astype(np.float32)
    # exact ties: the low 16 bits 0x8000, with even and odd bf16 mantissas
    ties = (np.arange(200, dtype=np.uint32) << 16 | 0x3F800000 | 0x8000).view(np.float32)
    a = np.concatenate([a, ties, -ties])
    want = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(M.bf16(a), want)


def test_quaternion_matches_scipy():
    R = Rotation.random(500, random_state=5)
    a = R.as_matrix().reshape(-1, 9)
    q = M.quaternion(a)
    s = R.as_quat()                      # (x, y, z, w)
    s = np.where(s[:, 3:4] < 0, -s, s)   # w >= 0
    keep = s[:, 3] > 1e-3
    np.testing.assert_allclose(q[keep], s[keep][:, [3, 0, 1, 2]], atol=1e-12)


def test_crystal_stress_closed_forms():
    rng = np.random.default_rng(1)
    lam = rng.normal(0, 50, (20, 6))
    eye = np.tile(np.eye(3).ravel(), (20, 1))
    np.testing.assert_allclose(M.crystal_stress(eye, lam), lam, atol=1e-12)
    a = Rotation.random(20, random_state=2).as_matrix().reshape(-1, 9)
    lc = M.crystal_stress(a, lam)
    np.testing.assert_allclose((lc * lc).sum(1), (lam * lam).sum(1), rtol=1e-12)
    np.testing.assert_allclose(lc[:, 5], lam[:, 5], atol=1e-12)       # the hydrostatic part is invariant
    # 90 degrees about z (crystal x -> sample y): a uniaxial sample-yy stress is crystal-xx
    az = Rotation.from_euler("z", 90, degrees=True).as_matrix().reshape(1, 9)
    syy = np.zeros((1, 6))
    syy[0] = np.einsum("ij,kij->k", np.diag([0.0, 1.0, 0.0]), M.BASIS)
    sc = M.crystal_stress(az, syy)
    np.testing.assert_allclose(np.einsum("k,kij->ij", sc[0], M.BASIS), np.diag([1.0, 0.0, 0.0]), atol=1e-12)


def test_start_rule_limits():
    u = np.tile(np.array([0, 1, 0, 0, 0, 0], float), (7, 1))
    y = np.array([[0, 0.5, 0, 0, 0, 0], [0, 10, 0, 0, 0, 0], [0, np.nan, 0, 0, 0, 0], [1, 1, 0, 0, 0, 0],
                  [0, 0, 0, 0, 0, 0], [3, -0.25, 0, 0, 0, 0], [0, -3, 0, 0, 0, 0]], float)
    nd = np.array([2.0, 2.0, 2.0, 0.0, 3.0, 2.0, 2.0])
    w = M.start(y, u, nd)
    np.testing.assert_allclose(w[0], [0, 2, 0, 0, 0, 0])     # never beyond the extrapolated step
    np.testing.assert_allclose(w[1], [0, 2, 0, 0, 0, 0])
    np.testing.assert_allclose(w[2], [0, 2, 0, 0, 0, 0])     # a non-finite correction: the extrapolated start
    np.testing.assert_array_equal(w[3], np.zeros(6))         # no previous change: the reference's start
    np.testing.assert_allclose(w[4], [0, 3, 0, 0, 0, 0])     # zero correction: the extrapolated start
    np.testing.assert_allclose(w[5], [0, 1.5, 0, 0, 0, 0])   # the component across d is not applied
    np.testing.assert_array_equal(w[6], np.zeros(6))          # a shortened step: down to the reference's start


@pytest.mark.parametrize("rounding", [True, False])
def test_features_layout(rounding):
    rng = np.random.default_rng(4)
    a = Rotation.random(8, random_state=4).as_matrix().reshape(-1, 9)
    lam = rng.normal(0, 30, (8, 6))
    d = rng.normal(0, 1e-3, (8, 6))
    d[0] = 0.0
    de = rng.normal(0, 1e-8, (8, 6))
    de[1] = 0.0
    rho = rng.uniform(0.5, 3, (8, 12))
    x, nd = M.features(a, lam, d, de, rho)
    assert x.shape == (8, 24) and x.dtype == np.float32
    np.testing.assert_allclose(np.linalg.norm(x[1:, 4:10], axis=1), 1.0, rtol=1e-6)
    np.testing.assert_array_equal(x[0, 4:10], 0.0)
    np.testing.assert_allclose(nd, np.linalg.norm(d, axis=1), rtol=1e-6)
    # the forcing is rotated like a stress: its norm is frame independent, and e / (|d| + |e|)
    en = M.E_SCALE * np.linalg.norm(de, axis=1)
    np.testing.assert_allclose(np.linalg.norm(x[:, 16:22], axis=1), en / (nd + en), rtol=1e-5)
    np.testing.assert_array_equal(x[1, 16:22], 0.0)
    np.testing.assert_allclose(np.linalg.norm(x[0, 16:22]), 1.0, rtol=1e-6)   # d = 0: a unit direction
    np.testing.assert_allclose(x[:, 23], np.log10(rho.mean(1)), rtol=1e-6)
    y = M.forward(M.random_blob(1), x, rounding=rounding)
    assert np.isfinite(y).all()
