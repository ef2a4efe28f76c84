"""CPU restatement of the learned Newton start (network, features, start rule).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU network (srx_mlp_forward,
srx_mlp_grad, the features the solver records).  The network has no counterpart in the
reference; what it must reproduce is the GPU path's definition (include/srx.h):

* forward: x' = (x - mean) * inv_spread (FP32), rounded to bf16; weights rounded to bf16;
  z1 = x' W1 + b1, h1 = bf16(relu(z1)); z2 = h1 W2 + b2, h2 = bf16(relu(z2)); y = h2 W3 + b3
  (products accumulated in FP32 on the GPU, in float64 here);
* loss: mean over samples and the 6 outputs of (y - (t - t_mean) / t_spread)^2;
* backward: straight-through bf16 (the rounded weights, inputs and hidden values of the forward
  pass, no derivative of the rounding).

Pinned by tests/test_mlp_oracle.py: the backward against central finite differences of the
unrounded float64 network, the bf16 rounding against torch's conversion, the quaternion against
scipy's, and the crystal-frame stress against an explicit tensor rotation.
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 24, 64, 6
O_W1 = 0
O_B1 = O_W1 + NIN * NH
O_W2 = O_B1 + NH
O_B2 = O_W2 + NH * NH
O_W3 = O_B2 + NH
O_B3 = O_W3 + NH * NOUT
NPARAM = O_B3 + NOUT
O_XM = NPARAM
O_XS = O_XM + NIN
O_TM = O_XS + NIN
O_TS = O_TM + NOUT
BLOB = O_TS + NOUT


def bf16(a) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(blob):
    b = np.asarray(blob)
    return dict(W1=b[O_W1:O_B1].reshape(NIN, NH), b1=b[O_B1:O_W2], W2=b[O_W2:O_B2].reshape(NH, NH),
                b2=b[O_B2:O_W3], W3=b[O_W3:O_B3].reshape(NH, NOUT), b3=b[O_B3:NPARAM], xm=b[O_XM:O_XS],
                xs=b[O_XS:O_TM], tm=b[O_TM:O_TS], ts=b[O_TS:BLOB])


def _forward(P, x, rounding):
    """rounding: the GPU path's float32 normalisation and bf16 rounding points; without it, the
    same network in float64 throughout."""
    r = bf16 if rounding else (lambda v: np.asarray(v, dtype=np.float64))
    if rounding:
        xn = ((np.asarray(x, np.float32) - P["xm"].astype(np.float32)).astype(np.float32) *
              P["xs"].astype(np.float32)).astype(np.float32)
    else:
        xn = (np.asarray(x, np.float64) - P["xm"].astype(np.float64)) * P["xs"].astype(np.float64)
    X = r(xn).astype(np.float64)
    W1, W2, W3 = (r(P[k]).astype(np.float64) for k in ("W1", "W2", "W3"))
    z1 = X @ W1 + P["b1"].astype(np.float64)
    h1 = r(np.maximum(z1, 0.0).astype(np.float32) if rounding else np.maximum(z1, 0.0)).astype(np.float64)
    z2 = h1 @ W2 + P["b2"].astype(np.float64)
    h2 = r(np.maximum(z2, 0.0).astype(np.float32) if rounding else np.maximum(z2, 0.0)).astype(np.float64)
    y = h2 @ W3 + P["b3"].astype(np.float64)
    return dict(X=X, W1=W1, W2=W2, W3=W3, z1=z1, h1=h1, z2=z2, h2=h2, y=y)


def forward(blob, x, rounding: bool = True) -> np.ndarray:
    """De-normalised outputs y * t_spread + t_mean for raw features x[n][24]."""
    P = unpack(blob)
    y = _forward(P, x, rounding)["y"]
    return y * P["ts"].astype(np.float64) + P["tm"].astype(np.float64)


def loss_grad(blob, x, t, rounding: bool = True):
    """(loss, gradient[NPARAM]) of the normalised mean squared error."""
    P = unpack(blob)
    F = _forward(P, x, rounding)
    tn = (np.asarray(t, np.float64) - P["tm"].astype(np.float64)) / P["ts"].astype(np.float64)
    e = F["y"] - tn
    n = x.shape[0]
    loss = float(np.mean(e * e))
    dy = 2.0 * e / (n * NOUT)
    g = np.zeros(NPARAM)
    g[O_W3:O_B3] = (F["h2"].T @ dy).ravel()
    g[O_B3:NPARAM] = dy.sum(0)
    dz2 = (dy @ F["W3"].T) * (F["z2"] > 0)
    g[O_W2:O_B2] = (F["h1"].T @ dz2).ravel()
    g[O_B2:O_W3] = dz2.sum(0)
    dz1 = (dz2 @ F["W2"].T) * (F["z1"] > 0)
    g[O_W1:O_B1] = (F["X"].T @ dz1).ravel()
    g[O_B1:O_W2] = dz1.sum(0)
    return loss, g


def quaternion(a9) -> np.ndarray:
    """(w, x, y, z), w >= 0, of orientation matrices a[n][9] (row-major)."""
    a = np.asarray(a9, np.float64).reshape(-1, 3, 3)
    w = 0.5 * np.sqrt(np.maximum(0.0, 1.0 + a[:, 0, 0] + a[:, 1, 1] + a[:, 2, 2]))
    x = np.copysign(0.5 * np.sqrt(np.maximum(0.0, 1.0 + a[:, 0, 0] - a[:, 1, 1] - a[:, 2, 2])), a[:, 2, 1] - a[:, 1, 2])
    y = np.copysign(0.5 * np.sqrt(np.maximum(0.0, 1.0 - a[:, 0, 0] + a[:, 1, 1] - a[:, 2, 2])), a[:, 0, 2] - a[:, 2, 0])
    z = np.copysign(0.5 * np.sqrt(np.maximum(0.0, 1.0 - a[:, 0, 0] - a[:, 1, 1] + a[:, 2, 2])), a[:, 1, 0] - a[:, 0, 1])
    return np.stack([w, x, y, z], 1)


_R2, _R3, _R6 = 1 / np.sqrt(2.0), 1 / np.sqrt(3.0), 1 / np.sqrt(6.0)
BASIS = np.zeros((6, 3, 3))
BASIS[0, 0, 0], BASIS[0, 1, 1] = -_R2, _R2
BASIS[1, 0, 0], BASIS[1, 1, 1], BASIS[1, 2, 2] = -_R6, -_R6, 2 * _R6
BASIS[2, 1, 2] = BASIS[2, 2, 1] = _R2
BASIS[3, 0, 2] = BASIS[3, 2, 0] = _R2
BASIS[4, 0, 1] = BASIS[4, 1, 0] = _R2
BASIS[5] = np.eye(3) * _R3


def crystal_stress(a9, lam6) -> np.ndarray:
    """B-basis components of a^T sigma a, sigma = sum_k lam6[k] B_k (a: crystal -> sample)."""
    a = np.asarray(a9, np.float64).reshape(-1, 3, 3)
    s = np.einsum("nk,kij->nij", np.asarray(lam6, np.float64), BASIS)
    sc = np.einsum("nji,njk,nkl->nil", a, s, a)
    return np.einsum("nij,kij->nk", sc, BASIS)


E_SCALE = 1.0e5   # MPa, MLP_E_SCALE of mlp_dev.hpp


def features(a9, lam6, d6, dec6, rho12) -> tuple[np.ndarray, np.ndarray]:
    """Raw features [n][24] and previous-change magnitudes [n] (definition in include/srx.h):
    quaternion, direction of d, lam / 100 MPa (crystal frame), the forcing e / (|d| + |e|) with
    e = E_SCALE * (eps_c - eps_c of the previous call) in the crystal frame (dec6 is that strain
    difference in the sample frame), log10(|d| / |lam|), log10 of the mean rho_s."""
    q = quaternion(a9)
    lc = crystal_stress(a9, lam6)
    ec = crystal_stress(a9, dec6)
    d = np.asarray(d6, np.float64)
    nd = np.sqrt((d * d).sum(1))
    u = np.where(nd[:, None] > 0, d / np.where(nd > 0, nd, 1.0)[:, None], 0.0)
    ln = np.sqrt((lc * lc).sum(1))
    en = E_SCALE * np.sqrt((ec * ec).sum(1))
    den = nd + en
    f = np.where(den[:, None] > 0, E_SCALE * ec / np.where(den > 0, den, 1.0)[:, None], 0.0)
    rho = np.asarray(rho12, np.float64).mean(1)
    x = np.concatenate([q, u, lc * 0.01, f, np.log10(np.maximum(nd, 1e-300) / np.maximum(ln, 1e-300))[:, None],
                        np.log10(np.maximum(rho, 1e-30))[:, None]], 1)
    return x.astype(np.float32), nd.astype(np.float32)


def start(y, u, nd) -> np.ndarray:
    """Newton start offsets [n][6]: alpha nd u, alpha = 1 + y . u (the network's de-normalised
    correction projected on the previous change's direction u) limited to [0, 1], 1 where not
    finite; zero where nd is zero or not finite."""
    y = np.asarray(y, np.float64)
    u = np.asarray(u, np.float64)
    with np.errstate(invalid="ignore"):
        alpha = 1.0 + (y * u).sum(1)
    alpha = np.where(np.isfinite(alpha), np.clip(alpha, 0.0, 1.0), 1.0)
    nd = np.asarray(nd, np.float64)
    ok = (nd > 0) & (nd < 1e30)
    return np.where(ok[:, None], (alpha * nd)[:, None] * u, 0.0)


def random_blob(seed: int = 0, xm=None, xs=None, tm=None, ts=None) -> np.ndarray:
    """A blob with He-uniform weights, small random biases and the given (or unit) statistics."""
    rng = np.random.default_rng(seed)
    b = np.zeros(BLOB, np.float32)
    for off, fan, cnt in ((O_W1, NIN, NIN * NH), (O_W2, NH, NH * NH), (O_W3, NH, NH * NOUT)):
        lim = np.sqrt(6.0 / fan)
        b[off:off + cnt] = rng.uniform(-lim, lim, cnt)
    b[O_B1:O_W2] = rng.normal(0, 0.1, NH)
    b[O_B2:O_W3] = rng.normal(0, 0.1, NH)
    b[O_B3:NPARAM] = rng.normal(0, 0.1, NOUT)
    b[O_XM:O_XS] = 0.0 if xm is None else xm
    b[O_XS:O_TM] = 1.0 if xs is None else xs
    b[O_TM:O_TS] = 0.0 if tm is None else tm
    b[O_TS:BLOB] = 1.0 if ts is None else ts
    return b
