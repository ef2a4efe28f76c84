"""The C-ABI library loads on a machine without a GPU, exports every symbol include/srx.h
declares, and returns the reference's setup error codes (no HIP call happens before them)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import srx_amd
from srx_amd import inputs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "srx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*(srx_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(srx_amd.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
    assert set(srx_amd.EXPORTS) <= set(names)


def test_strerror_and_default_options():
    L = srx_amd.lib()
    assert L.srx_strerror(1139).decode().startswith("error in reading the microstructure")
    assert L.srx_strerror(srx_amd.SrxError.__init__ and 3001).decode() == "NaN in the constitutive update"
    o = srx_amd.options()
    assert o.precision == 0 and o.fixed_iters == 0


def open_code(deck):
    L = srx_amd.lib()
    h = C.c_void_p()
    o = srx_amd.options()
    rc = L.srx_open(deck.encode(), os.path.dirname(deck).encode(), C.byref(o), C.byref(h))
    assert rc != 0, "expected a setup error"
    return rc


def test_setup_error_codes(tmp_path):
    inputs.make_case(str(tmp_path), n=(4, 4, 4), grains=2, seed=1, nsteps=1, timestep=0.1)
    d = str(tmp_path)
    bad = os.path.join(d, "bad116.in")
    inputs.write_deck(bad, n=(4, 4, 4), extra={"VelGrad_BC_Flag": "1 1 0 1 1 1", "Stress_BC_Flag": "0 0 1 0 0 0"})
    assert open_code(bad) == 116
    bad = os.path.join(d, "bad114.in")
    inputs.write_deck(bad, n=(4, 4, 4), extra={"Stress_BC_Flag": "1 1 1 0 0 0"})
    assert open_code(bad) == 114
    lines = open(os.path.join(d, "mesh.txt")).read().splitlines()
    lines[3], lines[4] = lines[4], lines[3]
    open(os.path.join(d, "mesh_swapped.txt"), "w").write("\n".join(lines) + "\n")
    bad = os.path.join(d, "bad1139.in")
    inputs.write_deck(bad, n=(4, 4, 4), mesh="mesh_swapped.txt")
    assert open_code(bad) == 1139
    assert open_code(os.path.join(d, "missing.in")) == 2001
    bad = os.path.join(d, "bad_grid.in")
    inputs.write_deck(bad, n=(7, 4, 4))
    assert open_code(bad) == 2002
    # every extent supported, but 2^28 voxels: beyond the 32-bit byte offsets of the field accesses
    bad = os.path.join(d, "bad_grid_size.in")
    inputs.write_deck(bad, n=(256, 1024, 1024))
    assert open_code(bad) == 2002
    bad = os.path.join(d, "bad_sx.in")
    inputs.write_deck(bad, n=(4, 4, 4), sx="nope.sx")
    assert open_code(bad) == 2001
    # slip system not orthogonal (init.c:195-197)
    sx = open(os.path.join(d, "Cu_p.sx")).read().replace("1 1 -1\t\t0 1 1", "1 1 -1\t\t1 1 1", 1)
    open(os.path.join(d, "bad.sx"), "w").write(sx)
    bad = os.path.join(d, "bad7.in")
    inputs.write_deck(bad, n=(4, 4, 4), sx="bad.sx")
    assert open_code(bad) == 7
    bad = os.path.join(d, "bad3002.in")
    inputs.write_deck(bad, n=(4, 4, 4), update_flag=0)
    assert open_code(bad) == 3002


def parse_ok(rc):
    """srx_open got past the mesh reader: 0 on a GPU, SRX_E_HIP (2003) where there is none."""
    return rc in (0, 2003)


def test_mesh_reader_line_forms(tmp_path, oracle_lib):
    """The mesh reader follows fgets(buf, 80) + sscanf (init.c:253-301) whether it parses in
    parallel (every line shorter than 79 characters) or record by record: CRLF line ends parse,
    an over-long line or a short file is error 1139, as in the oracle."""
    inputs.make_case(str(tmp_path), n=(4, 4, 4), grains=2, seed=3, nsteps=1, timestep=0.1)
    d = str(tmp_path)
    lines = open(os.path.join(d, "mesh.txt")).read().splitlines()
    variants = {
        "crlf": ("\r\n".join(lines) + "\r\n", True),
        "long": ("\n".join(lines[:5] + [lines[5] + " " * 90] + lines[6:]) + "\n", False),
        "short": ("\n".join(lines[:-1]) + "\n", False),
        "noeol": ("\n".join(lines), True),
    }
    for name, (text, ok) in variants.items():
        open(os.path.join(d, "mesh_%s.txt" % name), "w", newline="").write(text)
        deck = os.path.join(d, "deck_%s.in" % name)
        inputs.write_deck(deck, n=(4, 4, 4), mesh="mesh_%s.txt" % name)
        ro = oracle_lib.open_status(deck)
        rp = open_code(deck) if not ok else _open_any(deck)
        if ok:
            assert ro == 0 and parse_ok(rp), (name, ro, rp)
        else:
            assert ro == 1139 and rp == 1139, (name, ro, rp)


def _open_any(deck):
    L = srx_amd.lib()
    h = C.c_void_p()
    o = srx_amd.options()
    rc = L.srx_open(deck.encode(), os.path.dirname(deck).encode(), C.byref(o), C.byref(h))
    if rc == 0:
        L.srx_close(h)
    return rc


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
