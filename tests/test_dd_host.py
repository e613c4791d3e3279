"""Host set-up of the dislocation-density law (libevp: namelist keys,
Burgers lengths, sense vectors, interaction and projection tables) against
the oracle's, on the CPU."""
import ctypes as C
import os

import numpy as np

import oracle as O
from pyevp import lib, synth


def test_dd_host_tables_match_the_oracle(tmp_path):
    law = synth.DDLaw(inter=(0.11, 0.22, 0.33, 0.44, 0.55, 0.66))
    inp = synth.write_job(str(tmp_path), synth.Config(n=8, grains=4, increments=1, dd=law))
    J = O.Job(inp, O.LAW_DD)
    t = J.dd_tables()
    L = lib.evp()
    keys = (C.c_double * 64)()
    assert L.evp_parse_dd_keys(os.fsencode(inp), keys) == 0
    ph = lib.Phase()
    ph.is_gas, ph.nsys = 0, 12
    for j, (n, b) in enumerate(synth.FCC_SLIP):
        nn = np.array(n, float) / np.linalg.norm(n)
        bb = np.array(b, float) / np.linalg.norm(b)
        for m in range(3):
            ph.bn[j][m], ph.bb[j][m] = nn[m], bb[m]
    out = lib.DDPhase()
    L.evp_dd_shear_modulus.restype = C.c_double
    L.evp_dd_shear_modulus.argtypes = [C.POINTER(C.c_double)]
    cu = (C.c_double * 3)(241.1674e3, 89.1989e3, 151.9685e3)
    G = L.evp_dd_shear_modulus(cu)
    assert G == t["shear_G"][0]
    L.evp_dd_phase_setup.argtypes = [C.c_void_p, C.c_int, C.POINTER(lib.Phase), C.c_double,
                                     C.c_double, C.POINTER(lib.DDPhase)]
    L.evp_dd_phase_setup(keys, 0, C.byref(ph), G, 0.35, C.byref(out))
    np.testing.assert_array_equal(np.ctypeslib.as_array(out.chi), t["chi"][0])
    np.testing.assert_allclose(np.ctypeslib.as_array(out.cth), t["cth"][0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(np.ctypeslib.as_array(out.bp), t["bp"][0], rtol=0, atol=1e-15)
    np.testing.assert_array_equal(np.ctypeslib.as_array(out.bbB), t["bbB"][0])
    assert list(out.c) == list(law.c)
    assert (out.q_slip, out.q_bulk, out.poisson_nu) == (law.q_slip, law.q_bulk, 0.35)


def test_dd_keys_follow_the_namelist_rules(tmp_path):
    """The DD keys are read with the namelist's rules: the first definition of
    a key wins, '#' lines are comments, absent keys read as 0."""
    inp = tmp_path / "input.in"
    inp.write_text("# comment\nT__K\t300\nQ_slip\t0.2\t0.0\nT__K\t999\n#C_1 7\nC_1 0.48 0.5\n")
    L = lib.evp()
    keys = (C.c_double * 64)()
    assert L.evp_parse_dd_keys(os.fsencode(str(inp)), keys) == 0
    k = np.ctypeslib.as_array(keys)
    # layout of evp_dd_keys: a0[2], L0, GND_ScaleRatio, T__K, Q_slip[2], Q_bulk[2], C[8][2], ...
    assert k[4] == 300.0                     # T__K: the first definition
    assert (k[5], k[6]) == (0.2, 0.0)        # Q_slip
    assert (k[9], k[10]) == (0.48, 0.5)      # C_1 (the commented line skipped)
    assert k[0] == 0.0 and k[2] == 0.0       # a0, L0 absent
    assert L.evp_parse_dd_keys(os.fsencode(str(tmp_path / "missing.in")), keys) == -1
