"""Pins the oracle's tensor kinematics (kinematics.c) against closed forms and library routines."""
import os

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


# --- lattice reorientation (constitutive.c:142-273) and Schmid rotation (kinematics.c:430-463) ---

GOLD_SX = os.path.join(os.path.dirname(__file__), "golden", "reference_inputs", "Cu_p.sx")


def reference_slip_systems(path=GOLD_SX, mode=1):
    """(n, b) unit vectors of one mode of a .sx file, read from its text (the reference's
    Cu_p.sx), independently of the oracle's reader and of the package's FCC table."""
    lines = open(path).read().splitlines()
    start = next(i for i, l in enumerate(lines) if l.startswith("# Mode# %d " % mode))
    nsys = int(lines[start + 2].split()[0])
    hdr = next(i for i in range(start, len(lines)) if "Slip" in lines[i])
    rows = np.array([[float(v) for v in lines[hdr + 1 + k].split()] for k in range(nsys)])
    n = rows[:, :3] / np.linalg.norm(rows[:, :3], axis=1, keepdims=True)
    b = rows[:, 3:] / np.linalg.norm(rows[:, 3:], axis=1, keepdims=True)
    return n, b


def axial(w):
    """Axial vector of the skew part of w: w x = axial(w) cross x for skew w."""
    return np.array([w[2, 1], w[0, 2], w[1, 0]])


def orientation(tran, rot):
    t = np.ascontiguousarray(tran, dtype=float).ravel().copy()
    r = np.ascontiguousarray(rot, dtype=float).ravel().copy()
    lib().orc_orientation(dptr(t), dptr(r))
    return t.reshape(3, 3), r.reshape(3, 3)


def skew(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])


@pytest.mark.parametrize("angle", [2e-6, 1e-4, 1e-2, 0.3, 1.0, 2.5])
def test_orientation_is_the_exact_rotation_about_the_spin_axis(angle):
    """Orientation (constitutive.c:142-194) applies I + 2(W + W^2)/(1 + tan^2(|w|/2)) with
    W = skew(w tan(|w|/2)/|w|): the exact rotation by |w| about w (Rodrigues with the half-angle
    tangent), multiplied from the left.  scipy's rotation vector is the independent check; sin in
    place of tan, a flipped W or a right multiplication fail it."""
    rng = np.random.default_rng(int(angle * 1e6) + 7)
    for _ in range(10):
        t0 = Rotation.random(random_state=rng.integers(1 << 30)).as_matrix()
        axis = rng.standard_normal(3)
        w = axis / np.linalg.norm(axis) * angle
        t1, r = orientation(t0, skew(w))
        R = Rotation.from_rotvec(w).as_matrix()
        np.testing.assert_allclose(r, R, atol=2e-15)
        np.testing.assert_allclose(t1, R @ t0, atol=2e-15)
        np.testing.assert_allclose(t1 @ t1.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(t1) - 1.0) < 1e-14


def test_orientation_reads_only_the_axial_vector():
    """w is taken from rot[2][1], rot[0][2], rot[1][0] alone (constitutive.c:150-152): the other
    three off-diagonal entries and the diagonal are not read."""
    t0 = Rotation.random(random_state=5).as_matrix()
    w = np.array([0.2, -0.1, 0.3])
    rot = skew(w)
    rot[1, 2] = rot[2, 0] = rot[0, 1] = 123.0
    rot[0, 0] = rot[1, 1] = rot[2, 2] = -7.0
    t1, _ = orientation(t0, rot)
    np.testing.assert_allclose(t1, Rotation.from_rotvec(w).as_matrix() @ t0, atol=2e-15)


def test_orientation_small_spin_branch():
    """For |w| < 1e-6 the reference resets the norm to 1 but keeps tan(|w|/2)
    (constitutive.c:158-161), so it rotates by |w| tan(|w|/2) ~ |w|^2/2 about w, not by |w|."""
    t0 = Rotation.random(random_state=9).as_matrix()
    w = np.array([3e-7, -4e-7, 1.2e-7])
    t1, _ = orientation(t0, skew(w))
    wn = np.linalg.norm(w)
    R = Rotation.from_rotvec(w * np.tan(wn / 2)).as_matrix()
    np.testing.assert_allclose(t1, R @ t0, atol=1e-16)
    assert np.abs(t1 - Rotation.from_rotvec(w).as_matrix() @ t0).max() > 1e-8


@pytest.fixture(scope="module")
def rot_case(tmp_path_factory):
    """4x4x2 Voronoi case using the reference's Cu_p.sx (mode 1 slip systems), TimeStep 0.1."""
    import shutil
    from oracle.oracle import Oracle
    from srx_amd import inputs
    d = tmp_path_factory.mktemp("rot")
    deck = inputs.make_case(str(d), n=(4, 4, 2), grains=3, seed=4, nsteps=3, timestep=0.1)
    shutil.copy(GOLD_SX, os.path.join(str(d), "Cu_p.sx"))
    return deck


def _set_state(o, tran, velgrad, gamdot):
    o.set_field("tranmat", tran.reshape(o.nxyz, 9))
    o.set_field("velgrad", velgrad.reshape(o.nxyz, 9))
    o.set_field("gamdot", gamdot.reshape(o.nxyz, 12))


def test_update_orient_pure_spin_composes(rot_case):
    """update_orient (constitutive.c:196-273) with no slip and a uniform skew velocity gradient
    W: each call rotates every TranMat by dt*W from the left, so k calls give
    exp(k dt W) TranMat_0 (scipy) -- the sign, the time step and the order of the product."""
    from oracle.oracle import Oracle
    o = Oracle(rot_case)
    N = o.nxyz
    t0 = Rotation.random(N, random_state=3).as_matrix()
    omega = np.array([0.8, -0.5, 1.1])
    vg = np.broadcast_to(skew(omega), (N, 3, 3)).copy()
    _set_state(o, t0, vg, np.zeros((N, 12)))
    k = 5
    for _ in range(k):
        lib().orc_update_orient(o.h)
    t = o.field("tranmat").reshape(N, 3, 3)
    R = Rotation.from_rotvec(k * 0.1 * omega).as_matrix()
    np.testing.assert_allclose(t, R @ t0, atol=1e-14)
    np.testing.assert_allclose(np.einsum("pij,pkj->pik", t, t), np.broadcast_to(np.eye(3), (N, 3, 3)),
                               atol=1e-14)
    o.close()


def test_update_orient_pure_slip_leaves_the_lattice_unrotated(rot_case):
    """When the whole velocity gradient is carried by slip, L = sum_s gdot_s (T b_s) x (T n_s)
    with b, n from the reference's Cu_p.sx, the lattice spin skew(L) - skew(Lp) vanishes and
    TranMat is unchanged (constitutive.c:217-250).  A flipped plastic-spin sign, n x b in place
    of b x n, the transposed TranMat on b or n, or a misindexed slip rate all rotate it."""
    from oracle.oracle import Oracle
    o = Oracle(rot_case)
    N = o.nxyz
    n, b = reference_slip_systems()
    rng = np.random.default_rng(21)
    t0 = Rotation.random(N, random_state=8).as_matrix()
    g = rng.standard_normal((N, 12)) * 1e-2
    bs = np.einsum("pij,sj->psi", t0, b)
    ns = np.einsum("pij,sj->psi", t0, n)
    L = np.einsum("ps,psi,psj->pij", g, bs, ns)
    _set_state(o, t0, L, g)
    lib().orc_update_orient(o.h)
    np.testing.assert_allclose(o.field("tranmat").reshape(N, 3, 3), t0, atol=1e-16)
    # slip alone (L = 0): the lattice turns by -dt skew(Lp), i.e. about -dt axial(skew(Lp))
    _set_state(o, t0, np.zeros((N, 3, 3)), g)
    lib().orc_update_orient(o.h)
    t = o.field("tranmat").reshape(N, 3, 3)
    for p in range(N):
        wp = axial((L[p] - L[p].T) / 2)
        np.testing.assert_allclose(t[p], Rotation.from_rotvec(-0.1 * wp).as_matrix() @ t0[p], atol=1e-15)
    o.close()


def test_schmid_tensors_at_identity_are_the_slip_system_dyads(rot_case):
    """update_schmid (kinematics.c:430-463) at TranMat = I gives Schm_xt (init.c:200-206): the
    5 deviatoric B-basis components of sym(n x b) of the unit vectors of Cu_p.sx mode 1."""
    from oracle.oracle import Oracle
    o = Oracle(rot_case)
    N = o.nxyz
    o.set_field("tranmat", np.broadcast_to(np.eye(3).ravel(), (N, 9)).copy())
    lib().orc_update_schmid(o.h)
    n, b = reference_slip_systems()
    B = basis_closed_form()
    want = np.einsum("aij,sij->sa", B[:5], (np.einsum("si,sj->sij", n, b) + np.einsum("si,sj->sij", b, n)) / 2)
    out = np.zeros(60)
    for p in (0, N - 1):
        lib().orc_get_schmid(o.h, p, dptr(out))
        np.testing.assert_allclose(out.reshape(12, 5), want, atol=1e-15)
    o.close()


def test_schmid_tensors_resolve_the_rotated_shear_stress(rot_case):
    """For any TranMat T the rotated Schmid vector P_s resolves a stress as the sample-frame
    dyad does: P_s . sigma_B = (T b) . sigma . (T n) for symmetric sigma (B basis orthonormal).
    A transposed T, unnormalised b or n, or the rotation left out fail it."""
    from oracle.oracle import Oracle
    o = Oracle(rot_case)
    N = o.nxyz
    T = Rotation.random(N, random_state=12).as_matrix()
    o.set_field("tranmat", T.reshape(N, 9))
    lib().orc_update_schmid(o.h)
    n, b = reference_slip_systems()
    B = basis_closed_form()
    rng = np.random.default_rng(5)
    out = np.zeros(60)
    for p in range(N):
        lib().orc_get_schmid(o.h, p, dptr(out))
        P = out.reshape(12, 5)
        for _ in range(3):
            a = rng.standard_normal((3, 3))
            sig = a + a.T
            sb = np.einsum("aij,ij->a", B, sig)
            rss = np.einsum("si,ij,sj->s", (T[p] @ b.T).T, sig, (T[p] @ n.T).T)
            np.testing.assert_allclose(P @ sb[:5], rss, atol=1e-13)
    o.close()
