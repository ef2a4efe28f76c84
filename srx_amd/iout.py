"""Readers for the .iout field dumps and the reference's offline converters.

The solver writes, at the deck's PrintControl steps, the raw files of io.c:322-941
(`eps_<c>_S%06d.iout`, `els_…`, `sig_afterDRX_…`, `rho_afterDRX_SSD_S%06d.iout`,
`tex_afterDRX_S%06d.iout`, `slip_afterDRX_S%04d.iout`, `dislocation_afterDRX_S%04d.iout`,
`new_position_afterDRX_S%04d.iout`).  This module reads them into numpy and converts them to
the text the reference's converters print:

  tex2out.c      -> tex_to_text        "%lf\\t%lf\\t%lf\\t%d\\t%d\\t%d\\t%d\\t%d\\n"
  slip2out.c     -> sys_to_text        "%d" + 12 x "\\t%lf" + "\\n"   (slip and dislocation files)
  position2out.c -> position_to_text   "%lf\\t%lf\\t%lf\\n"
  out2iout.c     -> scalar_to_text ('b' direction, "%f\\n") and text_to_scalar ('f' direction)
  von_misces_ss.c -> von_mises_text    "%le\\n" of (1/sqrt2) sqrt((s11-s22)^2 + ... + 6 s12^2)

Command line: python -m srx_amd.iout {tex,slip,position,b,f,svm} <in> <out>
"""
from __future__ import annotations

import math
import sys

import numpy as np

# tex_afterDRX: {double angle[3]; int coord[3]; int jgr, jph, buff} (io.c:480-487)
TEX = np.dtype([("angle", "<f8", 3), ("coord", "<i4", 3), ("jgr", "<i4"), ("jph", "<i4"), ("buff", "<i4")])
# slip_afterDRX / dislocation_afterDRX: {int jgr; [4-byte hole]; double SS[12]} (io.c:632-640)
SYS = np.dtype({"names": ["jgr", "ss"], "formats": ["<i4", ("<f8", 12)], "offsets": [0, 8], "itemsize": 104})
# new_position_afterDRX: {double SS1, SS2, SS3} (io.c:725-733)
POS = np.dtype([("ss", "<f8", 3)])

SYM_ORDER = ("11", "22", "33", "23", "13", "12")   # component c = 1..6 of the tensor dumps


def read_scalar(path: str) -> np.ndarray:
    return np.fromfile(path, dtype="<f8")


def read_tex(path: str) -> np.ndarray:
    return np.fromfile(path, dtype=TEX)


def read_sys(path: str) -> np.ndarray:
    return np.fromfile(path, dtype=SYS)


def read_position(path: str) -> np.ndarray:
    return np.fromfile(path, dtype=POS)["ss"]


def read_tensor(stem: str, step: int, workdir: str = ".") -> np.ndarray:
    """The six files <stem>_<c>_S%06d.iout as an (N, 6) array in the order 11 22 33 23 13 12."""
    cols = [read_scalar("%s/%s_%01d_S%06d.iout" % (workdir, stem, c + 1, step)) for c in range(6)]
    return np.stack(cols, axis=1)


def tex_to_text(path_in: str, path_out: str) -> None:
    a = read_tex(path_in)
    with open(path_out, "w") as f:
        for r in a:
            f.write("%f\t%f\t%f\t%d\t%d\t%d\t%d\t%d\n" % (r["angle"][0], r["angle"][1], r["angle"][2],
                                                      r["coord"][0], r["coord"][1], r["coord"][2], r["jgr"], r["jph"]))


def sys_to_text(path_in: str, path_out: str) -> None:
    a = read_sys(path_in)
    with open(path_out, "w") as f:
        for r in a:
            f.write("%d\t" % r["jgr"] + "\t".join("%f" % x for x in r["ss"]) + "\n")


def position_to_text(path_in: str, path_out: str) -> None:
    a = read_position(path_in)
    with open(path_out, "w") as f:
        for r in a:
            f.write("%f\t%f\t%f\n" % (r[0], r[1], r[2]))


def scalar_to_text(path_in: str, path_out: str) -> None:
    a = read_scalar(path_in)
    with open(path_out, "w") as f:
        for x in a:
            f.write("%f\n" % x)


def text_to_scalar(path_in: str, path_out: str) -> None:
    with open(path_in) as f:
        vals = [float(t) for t in f.read().split()]
    np.asarray(vals, dtype="<f8").tofile(path_out)


def von_mises(sig6: np.ndarray) -> np.ndarray:
    """Per-voxel von Mises stress of von_misces_ss.c from (N, 6) components 11 22 33 23 13 12."""
    s = np.asarray(sig6, dtype=np.float64)
    d = ((s[:, 0] - s[:, 1]) ** 2 + (s[:, 0] - s[:, 2]) ** 2 + (s[:, 1] - s[:, 2]) ** 2
         + 6 * s[:, 3] ** 2 + 6 * s[:, 4] ** 2 + 6 * s[:, 5] ** 2)
    return (1 / math.sqrt(2)) * np.sqrt(d)


def von_mises_text(step: int, path_out: str, workdir: str = ".") -> None:
    v = von_mises(read_tensor("sig_afterDRX", step, workdir))
    with open(path_out, "w") as f:
        for x in v:
            f.write("%e\n" % x)


def main(argv) -> int:
    if len(argv) != 4:
        sys.stderr.write(__doc__)
        return 2
    kind, a, b = argv[1], argv[2], argv[3]
    if kind == "tex":
        tex_to_text(a, b)
    elif kind == "slip":
        sys_to_text(a, b)
    elif kind == "position":
        position_to_text(a, b)
    elif kind == "b":
        scalar_to_text(a, b)
    elif kind == "f":
        text_to_scalar(a, b)
    elif kind == "svm":
        von_mises_text(int(a), b)
    else:
        sys.stderr.write("unknown conversion %r\n" % kind)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
