"""Pins the oracle's tensor kinematics (kinematics.c) against closed forms and library routines."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle.oracle import dptr, lib

R2, R3, R6 = 1 / np.sqrt(2), 1 / np.sqrt(3), 1 / np.sqrt(6)


def basis_closed_form():
    """The Lebensohn-Tome basis of kinematics.c:35-53, written from its definition."""
    e = np.eye(3)
    B = np.zeros((6, 3, 3))
    B[0] = (-np.outer(e[0], e[0]) + np.outer(e[1], e[1])) * R2
    B[1] = (-np.outer(e[0], e[0]) - np.outer(e[1], e[1]) + 2 * np.outer(e[2], e[2])) * R6
    B[2] = (np.outer(e[1], e[2]) + np.outer(e[2], e[1])) * R2
    B[3] = (np.outer(e[0], e[2]) + np.outer(e[2], e[0])) * R2
    B[4] = (np.outer(e[0], e[1]) + np.outer(e[1], e[0])) * R2
    B[5] = np.eye(3) * R3
    return B


def v_to_t(v, dim=6):
    t = np.zeros(9)
    v = np.ascontiguousarray(v, dtype=float).copy()
    lib().orc_chg_basis(dptr(v), dptr(t), dptr(np.zeros(36)), dptr(np.zeros(81)), 1, dim)
    return t.reshape(3, 3)


def t_to_v(t, dim=6):
    v = np.zeros(6)
    t = np.ascontiguousarray(t, dtype=float).ravel().copy()
    lib().orc_chg_basis(dptr(v), dptr(t), dptr(np.zeros(36)), dptr(np.zeros(81)), 2, dim)
    return v[:dim]


def c4_to_c66(c4):
    c = np.zeros(36)
    lib().orc_chg_basis(dptr(np.zeros(6)), dptr(np.zeros(9)), dptr(c), dptr(np.ascontiguousarray(c4.ravel())), 4, 6)
    return c.reshape(6, 6)


def test_basis_matches_definition_and_is_orthonormal():
    B = basis_closed_form()
    for a in range(6):
        v = np.zeros(6)
        v[a] = 1.0
        np.testing.assert_allclose(v_to_t(v), B[a], atol=1e-15)
    G = np.einsum("aij,bij->ab", B, B)
    np.testing.assert_allclose(G, np.eye(6), atol=1e-15)


def test_roundtrip_symmetric_tensor():
    rng = np.random.default_rng(0)
    for _ in range(20):
        a = rng.standard_normal((3, 3))
        s = (a + a.T) / 2
        np.testing.assert_allclose(v_to_t(t_to_v(s)), s, atol=1e-14)
        # 5-dim variant drops the trace
        dev = s - np.trace(s) / 3 * np.eye(3)
        np.testing.assert_allclose(v_to_t(t_to_v(s, 5), 5), dev, atol=1e-14)


def iso_c4(lam, mu):
    d = np.eye(3)
    return (lam * np.einsum("ij,kl->ijkl", d, d) + mu * (np.einsum("ik,jl->ijkl", d, d) + np.einsum("il,jk->ijkl", d, d)))


def test_isotropic_stiffness_is_diagonal_in_basis():
    lam, mu = 70.0, 30.0
    c = c4_to_c66(iso_c4(lam, mu))
    np.testing.assert_allclose(c, np.diag([2 * mu] * 5 + [3 * lam + 2 * mu]), atol=1e-12)


def test_cubic_stiffness_is_diagonal_in_crystal_frame():
    """init.c:10-50 cubic C in the B basis: diag(c11-c12, c11-c12, 2c44, 2c44, 2c44, c11+2c12)."""
    c11, c12, c44 = 169665.0, 122500.0, 75827.0
    C = np.zeros((3, 3, 3, 3))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for l in range(3):
                    if i == j and k == l:
                        C[i, j, k, l] = c11 if i == k else c12
                    elif (i == k and j == l) or (i == l and j == k):
                        C[i, j, k, l] = c44
    np.testing.assert_allclose(c4_to_c66(C), np.diag([c11 - c12] * 2 + [2 * c44] * 3 + [c11 + 2 * c12]),
                               atol=1e-9)


def test_lu_inverse_matches_numpy():
    rng = np.random.default_rng(1)
    for _ in range(20):
        a = rng.standard_normal((6, 6)) + 6 * np.eye(6)
        c = a.copy().ravel()
        lib().orc_lu_inv66(dptr(c))
        np.testing.assert_allclose(c.reshape(6, 6), np.linalg.inv(a), rtol=1e-12, atol=1e-14)
    # the (commented out) self-test matrix of init.c:380-392
    s0 = np.zeros((6, 6))
    s0[0, 1] = 1.0
    s0[1, 0] = 2.0; s0[1, 2] = 2.0
    s0[2, 1] = 3.0; s0[2, 3] = 1.0
    s0[3, 2] = 1.0; s0[3, 4] = 2.0
    s0[4, 3] = 3.0; s0[4, 5] = 1.0
    s0[5, 4] = 2.0
    c = s0.copy().ravel()
    lib().orc_lu_inv66(dptr(c))
    np.testing.assert_allclose(c.reshape(6, 6) @ s0, np.eye(6), atol=1e-13)


def euler_matrix(t1, ph, t2):
    a = np.zeros(9)
    x = [np.array([t1]), np.array([ph]), np.array([t2])]
    lib().orc_euler_to_matrix(dptr(x[0]), dptr(x[1]), dptr(x[2]), dptr(a), 2)
    return a.reshape(3, 3)


def test_euler_matrix_is_bunge_zxz_passive():
    """EulerToTransMatrix opt 2 (kinematics.c:218-235) is the Bunge (ZXZ) sample->crystal
    matrix: the transpose of the active rotation Rz(phi1) Rx(Phi) Rz(phi2)."""
    rng = np.random.default_rng(2)
    for _ in range(20):
        t = rng.uniform([0, 0, 0], [2 * np.pi, np.pi, 2 * np.pi])
        a = euler_matrix(*t)
        r = Rotation.from_euler("ZXZ", t).as_matrix()
        np.testing.assert_allclose(a, r.T, atol=1e-14)
        np.testing.assert_allclose(a @ a.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(a) - 1) < 1e-14


def test_euler_roundtrip():
    rng = np.random.default_rng(3)
    for _ in range(20):
        t = rng.uniform([0, 0.2, 0], [np.pi, np.pi - 0.2, np.pi])
        a = euler_matrix(*t).ravel().copy()
        out = [np.zeros(1), np.zeros(1), np.zeros(1)]
        lib().orc_euler_to_matrix(dptr(out[0]), dptr(out[1]), dptr(out[2]), dptr(a), 1)
        deg = np.degrees(t)
        # opt 1 converts with the truncated PI of evp.h:84
        scale = 180.0 / 3.14159265359 / (180.0 / np.pi)
        got = np.array([out[0][0], out[1][0], out[2][0]])
        np.testing.assert_allclose(got, deg * scale, atol=1e-9)


@pytest.mark.parametrize("shape", [(4, 4, 4), (8, 2, 16), (3, 5, 6), (6, 6, 1), (12, 10, 9)])
def test_fft_matches_numpy(shape):
    rng = np.random.default_rng(4)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    buf = np.ascontiguousarray(x).view(np.float64).ravel().copy()
    lib().orc_fft3d(dptr(buf), *shape, -1, 0)
    np.testing.assert_allclose(buf.view(np.complex128).reshape(shape), np.fft.fftn(x), atol=1e-11)
    buf = np.ascontiguousarray(x).view(np.float64).ravel().copy()
    lib().orc_fft3d(dptr(buf), *shape, +1, 0)
    np.testing.assert_allclose(buf.view(np.complex128).reshape(shape), np.fft.ifftn(x) * x.size, atol=1e-11)


def green(c0, n, ks, dc=False, elast_bc=1, s0=None):
    du = np.zeros(18)
    s0 = np.zeros(81) if s0 is None else np.ascontiguousarray(s0.ravel())
    lib().orc_green_point(dptr(np.ascontiguousarray(c0.ravel())), dptr(np.ascontiguousarray(n, dtype=float)),
                          int(dc), elast_bc, dptr(s0), dptr(np.ascontiguousarray(ks).view(np.float64).ravel().copy()),
                          dptr(du))
    return du.view(np.complex128).reshape(3, 3)


def test_green_operator_equilibrium_and_isotropic_closed_form():
    """du = Gamma:sig makes sig + C0:sym(du) divergence free: (sig + C0:du).n = 0, and for an
    isotropic C0, Omega = (I - (lam+mu)/(lam+2mu) n n)/mu (init.c:542-592)."""
    rng = np.random.default_rng(5)
    lam, mu = 80.0, 40.0
    C0 = iso_c4(lam, mu)
    for _ in range(10):
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        s = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
        s = (s + s.T) / 2
        du = green(C0, n, s)
        t = s @ n
        om = (np.eye(3) - (lam + mu) / (lam + 2 * mu) * np.outer(n, n)) / mu
        np.testing.assert_allclose(du, -np.outer(om @ t, n), atol=1e-13)
        resid = (s + np.einsum("ijkl,kl->ij", C0, du)) @ n
        np.testing.assert_allclose(resid, 0, atol=1e-12)
    # anisotropic C0: equilibrium still holds
    A = rng.standard_normal((6, 6))
    c66 = A @ A.T + 6 * np.eye(6)
    c4 = np.zeros(81)
    lib().orc_chg_basis(dptr(np.zeros(6)), dptr(np.zeros(9)), dptr(c66.ravel().copy()), dptr(c4), 3, 6)
    c4 = c4.reshape(3, 3, 3, 3)
    n = np.array([0.3, -0.5, 0.81])
    n /= np.linalg.norm(n)
    s = rng.standard_normal((3, 3)) + 0j
    s = (s + s.T) / 2
    du = green(c4, n, s)
    np.testing.assert_allclose((s + np.einsum("ijkl,kl->ij", c4, du)) @ n, 0, atol=1e-11)
    # k = 0: Gamma = 0 (ElastBC 1) or -S0 (ElastBC 0)
    np.testing.assert_allclose(green(C0, np.zeros(3), s, dc=True), 0, atol=0)
    s0 = rng.standard_normal((3, 3, 3, 3))
    np.testing.assert_allclose(green(C0, np.zeros(3), s, dc=True, elast_bc=0, s0=s0),
                               -np.einsum("ijkl,kl->ij", s0, s), atol=1e-13)
