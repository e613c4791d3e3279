"""On-disk dumps of the drop-in (include/evp.h: evp_write_iout6,
evp_write_texture, evp_texture_to_text, evp_mat_to_euler), checked on the
CPU with synthetic fields: the byte layout src/io.c:283-541 writes and
src/tex2out.c:18-70 reads, and the matrix -> Bunge-angle inversion against
the oracle's EulerToTransMatrix(opt 2) restatement."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from pyevp import lib

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)
PI_REF = 3.14159265359

TEX = np.dtype([("angle", "<f8", (3,)), ("coord", "<i4", (3,)), ("grain", "<i4"),
                ("phase", "<i4"), ("buffer", "<i4")])


@pytest.fixture(scope="module")
def L():
    L = lib.evp()
    L.evp_mat_to_euler.argtypes = [DP, DP]
    L.evp_write_iout6.argtypes = [C.c_char_p, C.c_int, C.c_longlong, DP]
    L.evp_write_texture.argtypes = [C.c_char_p, C.c_int, IP, DP, IP, IP]
    L.evp_texture_to_text.argtypes = [C.c_char_p, C.c_char_p]
    return L


def p(a, t=DP):
    return a.ctypes.data_as(t)


def test_texture_record_is_48_bytes():
    assert TEX.itemsize == 48


@pytest.mark.parametrize("deg", [(30.0, 40.0, 50.0), (200.0, 100.0, 300.0), (10.0, 179.5, 20.0),
                                 (75.0, 0.2, 0.0)])
def test_mat_to_euler_inverts_the_oracle_rotation(L, deg):
    r = np.array(deg) * PI_REF / 180.0
    a = O.euler_to_mat(*r)
    ang = np.zeros(3)
    L.evp_mat_to_euler(p(np.ascontiguousarray(a).ravel()), p(ang))
    if abs(a[2, 2]) > 0.9999:
        # near the poles the reference treats the rotation as in-plane:
        # phi2 = 0, phi1 = atan2(a12, a11), Phi = acos(a33) -- lossy unless Phi
        # is exactly 0 or 180 degrees (src/kinematics.c:200-205)
        assert ang[2] == 0.0
        assert abs(ang[1] - np.degrees(np.arccos(a[2, 2])) * np.pi / PI_REF) < 1e-9
        assert abs(ang[0] - np.degrees(np.arctan2(a[0, 1], a[0, 0])) * np.pi / PI_REF) < 1e-9
        # exact at the pole itself: phi1 + phi2 survive as one in-plane angle
        a0 = O.euler_to_mat(0.3, 0.0, 0.4)
        L.evp_mat_to_euler(p(np.ascontiguousarray(a0).ravel()), p(ang))
        assert abs(ang[0] - 0.7 * 180.0 / PI_REF) < 1e-9 and ang[1] == 0.0 and ang[2] == 0.0
    else:
        t = np.array(deg)
        t[0], t[2] = (t[0] + 180.0) % 360.0 - 180.0, (t[2] + 180.0) % 360.0 - 180.0
        assert np.abs(ang - t).max() < 1e-9


def test_iout6_layout(L, tmp_path):
    n = 2 * 3 * 4
    f = np.arange(6 * n, dtype=np.float64).reshape(6, n) * 1.25 - 7.0
    pre = str(tmp_path / "sig").encode()
    assert L.evp_write_iout6(pre, 1500, n, p(f)) == 0
    for c in range(6):
        name = tmp_path / f"sig_{c + 1}_S001500.iout"
        assert name.stat().st_size == 8 * n
        assert np.array_equal(np.fromfile(name, dtype="<f8"), f[c])


def test_texture_layout_and_tex2out(L, tmp_path):
    dims = np.array([2, 3, 4], dtype=np.int32)
    n = int(np.prod(dims))
    rng = np.random.default_rng(5)
    deg = np.column_stack([rng.uniform(-170, 170, n), rng.uniform(10, 170, n), rng.uniform(-170, 170, n)])
    tran = np.zeros((9, n))
    for v in range(n):
        a = O.euler_to_mat(*(deg[v] * PI_REF / 180.0))
        tran[:, v] = a.T.ravel()  # TranMat_xt2sa = a^T
    grain = np.arange(n, dtype=np.int32) * 3 + 1
    phase = (np.arange(n, dtype=np.int32) % 2) + 1
    pre = str(tmp_path / "tex").encode()
    assert L.evp_write_texture(pre, 7, p(dims, IP), p(tran), p(grain, IP), p(phase, IP)) == 0
    path = tmp_path / "tex_S000007.iout"
    rec = np.fromfile(path, dtype=TEX)
    assert rec.shape == (n,)
    assert np.abs(rec["angle"] - deg).max() < 1e-9
    v = np.arange(n)
    assert np.array_equal(rec["coord"][:, 0], v // 12 + 1)
    assert np.array_equal(rec["coord"][:, 1], (v // 4) % 3 + 1)
    assert np.array_equal(rec["coord"][:, 2], v % 4 + 1)
    assert np.array_equal(rec["grain"], grain) and np.array_equal(rec["phase"], phase)
    assert np.all(rec["buffer"] == 0)
    txt = tmp_path / "tex.txt"
    exe = os.path.join(lib.LIBDIR, "tex2out")
    subprocess.run([exe, str(path), str(txt)], check=True)
    rows = np.loadtxt(txt)
    assert rows.shape == (n, 8)
    assert np.abs(rows[:, :3] - deg).max() < 1e-5  # "%lf": 6 decimals
    assert np.array_equal(rows[:, 6].astype(int), grain)
