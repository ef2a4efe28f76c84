"""Record layouts of the .iout dumps and the text of the offline converters (tex2out.c,
slip2out.c, position2out.c, out2iout.c, von_misces_ss.c), checked against C's own printf."""
import ctypes
import math

import numpy as np

from srx_amd import iout

libc = ctypes.CDLL(None)


def cfmt(fmt, *args):
    buf = ctypes.create_string_buffer(4096)
    cargs = [ctypes.c_double(a) if isinstance(a, float) else ctypes.c_int(a) for a in args]
    libc.snprintf(buf, 4096, fmt.encode(), *cargs)
    return buf.value.decode()


def test_record_sizes_match_the_c_structs():
    assert iout.TEX.itemsize == 48
    assert iout.SYS.itemsize == 104 and iout.SYS.fields["ss"][1] == 8
    assert iout.POS.itemsize == 24


def test_tex_to_text_matches_c_printf(tmp_path):
    rng = np.random.default_rng(0)
    a = np.zeros(5, dtype=iout.TEX)
    a["angle"] = rng.uniform(-180, 180, (5, 3))
    a["coord"] = rng.integers(1, 64, (5, 3))
    a["jgr"] = rng.integers(1, 1000, 5)
    a["jph"] = 1
    a.tofile(tmp_path / "t.iout")
    iout.tex_to_text(str(tmp_path / "t.iout"), str(tmp_path / "t.txt"))
    want = "".join(cfmt("%lf\t%lf\t%lf\t%d\t%d\t%d\t%d\t%d\n", *map(float, r["angle"]), *map(int, r["coord"]),
                        int(r["jgr"]), int(r["jph"])) for r in a)
    assert open(tmp_path / "t.txt").read() == want


def test_slip_and_position_to_text_match_c_printf(tmp_path):
    rng = np.random.default_rng(1)
    s = np.zeros(4, dtype=iout.SYS)
    s["jgr"] = [3, 7, 11, 12]
    s["ss"] = rng.standard_normal((4, 12)) * 1e-3
    s.tofile(tmp_path / "s.iout")
    raw = np.fromfile(tmp_path / "s.iout", dtype=np.uint8).reshape(4, 104)
    assert (raw[:, 4:8] == 0).all()
    iout.sys_to_text(str(tmp_path / "s.iout"), str(tmp_path / "s.txt"))
    fmt = "%d" + "\t%lf" * 12 + "\n"
    want = "".join(cfmt(fmt, int(r["jgr"]), *map(float, r["ss"])) for r in s)
    assert open(tmp_path / "s.txt").read() == want
    p = rng.standard_normal((6, 3))
    p.astype("<f8").tofile(tmp_path / "p.iout")
    iout.position_to_text(str(tmp_path / "p.iout"), str(tmp_path / "p.txt"))
    assert open(tmp_path / "p.txt").read() == "".join(cfmt("%lf\t%lf\t%lf\n", *map(float, r)) for r in p)


def test_out2iout_round_trip(tmp_path):
    x = np.array([0.0, -1.5, 3.25, 1e-7, 12345.678901])
    x.tofile(tmp_path / "x.iout")
    iout.scalar_to_text(str(tmp_path / "x.iout"), str(tmp_path / "x.txt"))
    assert open(tmp_path / "x.txt").read() == "".join(cfmt("%f\n", float(v)) for v in x)
    iout.text_to_scalar(str(tmp_path / "x.txt"), str(tmp_path / "y.iout"))
    np.testing.assert_array_equal(iout.read_scalar(str(tmp_path / "y.iout")), np.round(x, 6))


def test_von_mises_closed_forms():
    """Uniaxial s33 gives |s33|; pure shear s12 gives sqrt(3)|s12|; hydrostatic gives 0."""
    s = np.zeros((3, 6))
    s[0, 2] = -7.0
    s[1, 5] = 2.0
    s[2, :3] = 5.0
    np.testing.assert_allclose(iout.von_mises(s), [7.0, 2.0 * math.sqrt(3.0), 0.0], atol=1e-14)
