"""CPU checks of the product's boundary: the libraries load, export every symbol
include/*.h declares, and the C drop-in parses the reference's own input files."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from pyevp import lib

HEADERS = {"evpgpu.h": "libevpgpu.so", "evp.h": "libevp.so"}


def declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(\w+)\s*\([^;{]*\)\s*;", src, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_library_exports_every_declared_symbol(header):
    so = os.path.join(lib.LIBDIR, HEADERS[header])
    assert os.path.exists(so), f"{so} not built"
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True,
                         check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    names = declared(header)
    assert len(names) >= 7
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_python_symbol_lists_match_headers():
    assert sorted(lib.EVPGPU_SYMBOLS) == declared("evpgpu.h")
    assert sorted(lib.EVP_SYMBOLS) == declared("evp.h")


def test_libraries_load_without_gpu():
    g = lib.gpu()
    assert g.evpgpu_strerror(1203).decode().startswith("roundoff")
    e = lib.evp()
    for n in lib.EVP_SYMBOLS:
        assert hasattr(e, n)


def test_create_reports_status_instead_of_aborting():
    h = C.c_void_p()
    # bad sizes (each length 2..512) are rejected before any device call
    assert lib.gpu().evpgpu_create(C.byref(h), 1, 8, 8, 0) == -3
    assert lib.gpu().evpgpu_create(C.byref(h), 8, 513, 8, 0) == -3
    assert lib.gpu().evpgpu_create(C.byref(h), 8, 8, 8, 7) == -1


def test_namelist_and_printnamelist_on_reference_input(capfd):
    """GetNameList/PrintNameList of the drop-in on test_run/input.in."""
    L = lib.evp()
    argv = (C.c_char_p * 2)(b"forge", os.fsencode(os.path.join(GOLDEN, "test_run", "input.in")))
    L.GetNameList(argv)
    L.PrintNameList()
    out = capfd.readouterr().out
    assert "CellDim\t\t32 32 32" in out
    assert "TimeStep\t5.000000e-01" in out
    assert "VelGrad_BC_Flag\t0 0 1 1 1 1" in out
    assert "VelGrad_BC\t-5.000000e-04 -5.000000e-04 1.000000e-03 0.000000e+00 0.000000e+00 0.000000e+00" in out
    assert "Err\t\t5.000000e-05" in out
    assert "initial_ms\tFFT_microstructure_pore_197.txt" in out
    assert "** no data" not in out


class PhaseOut(lib.Phase):
    pass


def test_sx_parser_on_reference_file():
    """PlasticInit semantics on test_run/Cu_p.sx: mode 1 selected out of 3; the
    hardening matrix comes from the last mode listed (Q9)."""
    L = lib.evp()
    L.evp_parse_sx = getattr(C.CDLL(os.path.join(lib.LIBDIR, "libevp.so")), "evp_parse_sx")
    L.evp_parse_sx.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(lib.Phase)]
    ph = lib.Phase()
    st = L.evp_parse_sx(os.fsencode(os.path.join(GOLDEN, "test_run")), b"Cu_p.sx", C.byref(ph))
    assert st == 0
    assert ph.nsys == 12
    assert ph.nrs[0] == 10.0 and ph.gam0[0] == 1.0
    assert list(ph.tau[0]) == [150.0, 150.0, 5.0]
    assert list(ph.thet[0]) == [400.0, 250.0]
    bn = np.array([ph.bn[s][:] for s in range(12)])
    bb = np.array([ph.bb[s][:] for s in range(12)])
    assert np.allclose(np.linalg.norm(bn, axis=1), 1.0, atol=1e-15)
    assert np.allclose(np.sum(bn * bb, axis=1), 0.0, atol=1e-15)
    # normals of the first three systems: (1,1,-1)/sqrt(3)
    assert np.allclose(bn[0], np.array([1, 1, -1]) / np.sqrt(3), atol=1e-16)
    H = np.array([ph.hard[i][:12] for i in range(12)])
    assert np.all(H == 1.0)


def test_forge_binary_exists_and_requires_argument():
    forge = os.path.join(lib.LIBDIR, "forge")
    assert os.access(forge, os.X_OK)
    r = subprocess.run([forge], capture_output=True, text=True)
    assert r.returncode != 0
