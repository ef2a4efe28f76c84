"""Input files for the solver, in the reference's own formats.

* ``write_deck``  -- a namelist input deck (io.c:25-94 key set) for the common
  uniaxial-tension workload: x3 tension at a given rate with stress-free
  lateral faces (VelGrad_BC_Flag 0 0 1 1 1 1, Stress_BC_Flag 1 1 0 0 0 0),
  Cu with phase 2 as gas.
* ``write_sx``    -- a slip-mode file in the layout PlasticInit reads
  (init.c:52-160) with the 12 FCC {111}<110> systems.  The system order is
  the one of the reference's Cu_p.sx mode 1, because the Arsenlis-Parks
  interaction table (init.c:594-703) is indexed by that order.
* ``voronoi``     -- periodic Voronoi polycrystals: seeds uniform in the box,
  voxel centres assigned to the nearest seed under the minimum-image
  distance, orientations uniform (Shoemake quaternions) or a cube texture.
* ``write_mesh``  -- the 8-column voxel file InitMicroStruct reads
  (init.c:253-301): ``phi1 Phi phi2 i j k grain phase``, degrees, 1-based,
  x slowest, every line shorter than 80 characters.
"""
from __future__ import annotations

import os
import numpy as np

# (n, b) of the 12 {111}<110> systems, Cu_p.sx mode 1 order.
FCC_SLIP = [
    ((1, 1, -1), (0, 1, 1)), ((1, 1, -1), (1, 0, 1)), ((1, 1, -1), (1, -1, 0)),
    ((1, -1, -1), (0, 1, -1)), ((1, -1, -1), (1, 0, 1)), ((1, -1, -1), (1, 1, 0)),
    ((1, -1, 1), (0, 1, 1)), ((1, -1, 1), (1, 0, -1)), ((1, -1, 1), (1, 1, 0)),
    ((1, 1, 1), (0, 1, -1)), ((1, 1, 1), (1, 0, -1)), ((1, 1, 1), (1, -1, 0)),
]

PI_REF = 3.14159265359  # evp.h:84, used by the reader for deg->rad


def write_sx(path: str, tau0: float = 0.1, nsys: int = 12) -> None:
    """Slip-mode file with one FCC mode (label 1), tau0xf = ``tau0`` MPa: the first ``nsys`` of the
    12 {111}<110> systems in Cu_p.sx's order."""
    if not 1 <= nsys <= 12:
        raise ValueError("nsys must be 1..12")
    lines = [
        "# Crystal system (XtalSys):", "CUBIC", "",
        "# Crystal axex (XtalAxis):", "1.0 1.0 1.0", "",
        "# Total number of modes listed in the file (N_modes_max):", "1", "",
        "# Number of modes to be used in the calculation (N_modes):", "1", "",
        "# Labels of the modes to be used (iMode):", "1", "",
        "# Mode# 1 -- <110>(111) SLIP",
        "\tnsmx\tnrsx\tgamd0x\ttwshx\tisectwx",
        "\t%d\t\t10.\t\t1.0\t\t0.0\t\t0" % nsys,
        "\ttau0xf\ttau0xb\ttau1x\tthet0\tthet1",
        f"\t{tau0!r}\t\t{tau0!r}\t\t5.\t\t400.\t250.",
        "\thselfx\thlatex",
        "\t1.0\t\t1.0",
        "\tSlip (n-b)",
    ]
    for n, b in FCC_SLIP[:nsys]:
        lines.append("\t%d %d %d\t\t%d %d %d" % (n + b))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_deck(path: str, n=(16, 16, 16), mesh: str = "mesh.txt", sx: str = "Cu_p.sx",
               nsteps: int = 20, timestep: float = 1.0, rate: float = 1.0e-3,
               err: float = 2e-5, itermax: int = 1000, rho0: float = 1.0,
               print_control=(0, 50), update_flag: int = 1, hard_flag: int = 1,
               elast_bc: int = 1, extra: dict | None = None) -> None:
    """Common deck: x3 tension at ``rate`` 1/s, lateral faces stress free."""
    kv = [
        ("N_phases", "2"), ("CellDim", "%d %d %d" % tuple(n)), ("initial_ms", mesh),
        ("Type_phases", "0 1"), ("ElastConst_PhaseI", "162700. 119000. 71150. 0.35"),
        ("Slip_PhaseI", sx), ("ElastConst_PhaseII", "0.0 0.0 0.0 0.0"), ("Slip_PhaseII", "XXXXX"),
        ("VelGrad_BC_Flag", "0 0 1 1 1 1"), ("VelGrad_BC", "0.0 0.0 %r 0.0 0.0 0.0" % rate),
        ("Stress_BC_Flag", "1 1 0 0 0 0"), ("Stress_BC", "0. 0. 0. 0. 0. 0."),
        ("ElastBC", str(elast_bc)), ("TimeStep", repr(timestep)), ("N_steps", str(nsteps)),
        ("Err", repr(err)), ("IterMax", str(itermax)),
        ("PrintControl", "%d %d" % tuple(print_control)), ("Update_Flag", str(update_flag)),
        ("Hard_Flag", str(hard_flag)), ("Tex_Flag", "1"),
        ("a0", "0.362 0.362"), ("L0", "0.0045"), ("GND_ScaleRatio", "200.0"),
        ("Q_slip", "2.0 2.0"), ("Q_bulk", "1.6 1.6"), ("rho_SSD_initial", repr(rho0)),
        ("T__K", "300"), ("C_1", "1.0 1.0"), ("C_2", "1.0 1.0"), ("C_3", "1.0 1.0"),
        ("C_4", "600.0 600.0"), ("C_5", "17.6 17.6"), ("C_6", "0E0 0.0"), ("C_7", "0.0E-9 0.0"),
        ("C_8", "0.24 0.24"), ("Selfinter0", "0.2 0.2"), ("Coplanar0", "0.2 0.2"),
        ("CrossSlip0", "0.2 0.2"), ("GlissileJunction0", "0.3 0.3"), ("HirthLock0", "0.2 0.2"),
        ("LomerCottrellLock0", "0.4 0.4"), ("D_0", "1.5E-3"), ("nucleus_radius", "4"),
        ("CellDim_pf", "%d %d %d" % tuple(n)), ("RandSeed", "123"), ("E_gb", "0.625"),
        ("M_gb", "7.4E-11"), ("Scale_Fdeform", "0.5"), ("k_c", "320000.0"), ("alpha", "8.5"),
        ("zeta_kc", "1.0"), ("age_drx", "1E0"), ("ssdScaler_DRX", "0.74"), ("gndScaler_DRX", "0.66"),
        ("sigScaler_DRX", "0.66"), ("Nucl_Static_Flag", "1"), ("tau_DRX", "1000.0"),
        ("DeltaQ_DRX", "0.0"), ("DeltaTau_DRX", "0.0"), ("pf_length_scale", "2.5E-06"),
        ("pf_time_step", "0.004"),
    ]
    if extra:
        d = dict(kv)
        d.update({k: str(v) for k, v in extra.items()})
        kv = list(d.items())
    with open(path, "w") as f:
        f.write("# EVP-FFT deck: x3 uniaxial tension, lateral faces stress free\n")
        for k, v in kv:
            f.write("%s\t%s\n" % (k, v))
        f.write("\n")


def random_rotations(rng: np.random.Generator, count: int) -> np.ndarray:
    """Uniform random rotation matrices (Shoemake's unit quaternions)."""
    u1, u2, u3 = rng.random(count), rng.random(count), rng.random(count)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    return quat_to_matrix(q)


def quat_to_matrix(q: np.ndarray) -> np.ndarray:
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    r = np.empty((len(q), 3, 3))
    r[:, 0, 0] = 1 - 2 * (y * y + z * z); r[:, 0, 1] = 2 * (x * y - z * w); r[:, 0, 2] = 2 * (x * z + y * w)
    r[:, 1, 0] = 2 * (x * y + z * w); r[:, 1, 1] = 1 - 2 * (x * x + z * z); r[:, 1, 2] = 2 * (y * z - x * w)
    r[:, 2, 0] = 2 * (x * z - y * w); r[:, 2, 1] = 2 * (y * z + x * w); r[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return r


def cube_texture(rng: np.random.Generator, count: int, sigma_deg: float = 10.0) -> np.ndarray:
    """{100}<001> cube orientation spread by a Gaussian misorientation."""
    axis = rng.normal(size=(count, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    th = np.deg2rad(rng.normal(0.0, sigma_deg, count))
    q = np.concatenate([axis * np.sin(th / 2)[:, None], np.cos(th / 2)[:, None]], axis=1)
    return quat_to_matrix(q)


def matrix_to_bunge(a: np.ndarray) -> np.ndarray:
    """Bunge angles in degrees of sample->crystal matrices, the inverse of
    EulerToTransMatrix opt 2 as its opt 1 states it (kinematics.c:198-217)."""
    a = np.asarray(a, dtype=float)
    ph = np.arccos(np.clip(a[:, 2, 2], -1.0, 1.0))
    sph = np.sin(ph)
    special = np.abs(a[:, 2, 2]) > 0.9999
    with np.errstate(divide="ignore", invalid="ignore"):
        t1 = np.where(special, np.arctan2(a[:, 0, 1], a[:, 0, 0]), np.arctan2(a[:, 2, 0] / sph, -a[:, 2, 1] / sph))
        t2 = np.where(special, 0.0, np.arctan2(a[:, 0, 2] / sph, a[:, 1, 2] / sph))
    return np.degrees(np.stack([t1, ph, t2], axis=1))


def voronoi_grains(n, grains: int, seed: int, texture: str = "random", sigma_deg: float = 10.0):
    """Periodic Voronoi polycrystal on an n0 x n1 x n2 grid.

    Returns (euler_deg[grains,3], grain_id[N]): the Bunge angles of each grain and the grain id
    (1..grains) of every voxel in voxel order (x slowest)."""
    from scipy.spatial import cKDTree
    n = tuple(int(v) for v in n)
    rng = np.random.default_rng(seed)
    pts = rng.random((grains, 3)) * np.array(n, dtype=float)
    tree = cKDTree(pts, boxsize=np.array(n, dtype=float))
    gx, gy, gz = np.meshgrid(np.arange(n[0]) + 0.5, np.arange(n[1]) + 0.5, np.arange(n[2]) + 0.5, indexing="ij")
    centres = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)
    _, owner = tree.query(centres, k=1, workers=-1)
    rot = random_rotations(rng, grains) if texture == "random" else cube_texture(rng, grains, sigma_deg)
    # matrix_to_bunge expects sample->crystal matrices; the transpose of a uniform rotation is uniform.
    eul = matrix_to_bunge(np.transpose(rot, (0, 2, 1)))
    return eul, owner.astype(np.int64) + 1


def voronoi(n, grains: int, seed: int, texture: str = "random", sigma_deg: float = 10.0):
    """Periodic Voronoi polycrystal on an n0 x n1 x n2 grid.

    Returns (euler_deg[N,3], grain_id[N]) in voxel order (x slowest), with
    grain ids 1..grains."""
    eul, gid = voronoi_grains(n, grains, seed, texture, sigma_deg)
    return eul[gid - 1], gid


def write_mesh(path: str, n, euler_deg: np.ndarray, grain: np.ndarray, phase=None) -> None:
    n = tuple(int(v) for v in n)
    N = n[0] * n[1] * n[2]
    if phase is None:
        phase = np.ones(N, dtype=np.int64)
    idx = np.arange(N)
    i = idx // (n[1] * n[2]) + 1
    j = (idx // n[2]) % n[1] + 1
    k = idx % n[2] + 1
    e = np.asarray(euler_deg, dtype=float)
    with open(path, "w") as f:
        chunk = 1 << 16
        for s in range(0, N, chunk):
            sl = slice(s, min(N, s + chunk))
            f.write("".join("%.6f %.6f %.6f %d %d %d %d %d\n" % row for row in zip(
                e[sl, 0], e[sl, 1], e[sl, 2], i[sl], j[sl], k[sl], grain[sl], phase[sl])))


def _digits(width: int, count: int) -> np.ndarray:
    """Right-aligned decimal text of 0..count-1, `width` bytes each."""
    v = np.arange(count)
    out = np.full((count, width), ord(" "), dtype=np.uint8)
    rem = v.copy()
    for c in range(width - 1, -1, -1):
        d = (rem % 10).astype(np.uint8) + ord("0")
        out[:, c] = np.where((rem > 0) | (c == width - 1), d, ord(" "))
        rem //= 10
    return out


def write_mesh_grains(path: str, n, euler_grain: np.ndarray, grain: np.ndarray, phase: int = 1) -> None:
    """The same 8-column records as write_mesh for a mesh whose angles are those of its grain,
    written as fixed-width fields (%11.6f angles, %4d indices, %5d grain): 59-character lines,
    built with array operations (16.7 M lines in a few seconds).  Extents up to 9999 and up to
    99999 grains."""
    n = tuple(int(v) for v in n)
    N = n[0] * n[1] * n[2]
    G = len(euler_grain)
    if max(n) > 9999 or G > 99999 or not 0 <= phase <= 9:
        raise ValueError("write_mesh_grains: extents <= 9999, grains <= 99999, one-digit phase")
    ang = np.frombuffer("".join("%11.6f %11.6f %11.6f " % tuple(r) for r in euler_grain).encode(),
                        dtype=np.uint8).reshape(G, 36)
    d4 = _digits(4, max(n) + 1)
    d5 = _digits(5, G + 1)
    grain = np.asarray(grain, dtype=np.int64)
    L = 59
    with open(path, "wb") as f:
        chunk = 1 << 20
        for s in range(0, N, chunk):
            v = np.arange(s, min(N, s + chunk))
            buf = np.full((len(v), L), ord(" "), dtype=np.uint8)
            buf[:, 0:36] = ang[grain[v] - 1]
            buf[:, 36:40] = d4[v // (n[1] * n[2]) + 1]
            buf[:, 41:45] = d4[(v // n[2]) % n[1] + 1]
            buf[:, 46:50] = d4[v % n[2] + 1]
            buf[:, 51:56] = d5[grain[v]]
            buf[:, 57] = ord("0") + phase
            buf[:, 58] = ord("\n")
            buf.tofile(f)


def make_case(workdir: str, n=(16, 16, 16), grains: int = 8, seed: int = 1, nsteps: int = 20,
              texture: str = "random", mesh_name: str = "mesh.txt", deck_name: str = "deck.in",
              nsys: int = 12, **deck_kw) -> str:
    """Write sx + mesh + deck for a Voronoi case into ``workdir``; return the deck path."""
    os.makedirs(workdir, exist_ok=True)
    write_sx(os.path.join(workdir, "Cu_p.sx"), nsys=nsys)
    eul, gid = voronoi_grains(n, grains, seed, texture)
    write_mesh_grains(os.path.join(workdir, mesh_name), n, eul, gid)
    deck = os.path.join(workdir, deck_name)
    write_deck(deck, n=n, mesh=mesh_name, nsteps=nsteps, **deck_kw)
    return deck
