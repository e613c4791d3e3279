"""The Green operator at one wave vector of the half spectrum
(m-themes_amd/csrc/green_point.h, the code the x-line kernel and the Green
pass run; host build in tests/hostcheck) against the oracle's operator on
the full grid (oracle/evp_oracle.c green_point: Gamma from src/init.c:541-591
applied as src/evolution.c:1373-1387).

Away from the split Nyquist entries the half-spectrum result is the oracle's
at the same wave vector.  At a split entry the library applies
Gamma_eff = (Gamma(n(k)) + Gamma(n(-k)))/2, so the oracle is run a second
time with the same stress transform placed at -k and the two results are
averaged.  The symmetric part (11 22 33 23 13 12) and the antisymmetric part
(23 13 12) are compared; FP64 to 1e-13 of the largest entry (measured
1.3e-15), FP32 to 2e-6 (measured 4.5e-7).  Without the averaging the split
entries differ at order one, so the Gamma_eff rule is what is checked there."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hostcheck")
LIB = os.path.join(HERE, "lib", "libhostcheck.so")
COMP = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
ANTI = [(1, 2), (0, 2), (0, 1)]


@pytest.fixture(scope="module")
def hostlib():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    H = C.CDLL(LIB)
    H.gpc_green_point.restype = C.c_int
    H.gpc_green_point.argtypes = [C.c_int, O.DP] + [C.c_int] * 6 + [O.DP, O.DP, O.DP]
    return H


def green_coefficients(C0):
    """A = C0 : (n x n) as 36 coefficients of (n0^2, n1^2, n2^2, n1n2, n0n2,
    n0n1), the layout the library's Green parameters use."""
    ga = np.zeros(36)
    for q, (i, k) in enumerate(COMP):
        for p, (j, l) in enumerate(COMP):
            ga[q * 6 + p] = C0[i, j, k, l] if j == l else C0[i, j, k, l] + C0[i, l, k, j]
    return ga


@pytest.mark.parametrize("dims", [(8, 6, 10), (6, 5, 4)])
def test_half_spectrum_operator_vs_oracle(job_factory, hostlib, dims):
    nx, ny, nz = dims
    J = O.Job(job_factory(f"gp{nx}{ny}{nz}", n=0, dims=dims, grains=4, seed=9, increments=1))
    _, _, C0, _ = J.setup_info()
    ga = green_coefficients(C0)
    rng = np.random.default_rng(3)
    a = rng.standard_normal((nx, ny, nz, 3, 3)) + 1j * rng.standard_normal((nx, ny, nz, 3, 3))
    s = (a + np.swapaxes(a, -1, -2)) * 50.0
    # the same stress transform at -k: s_neg[-k] = s[k]
    ix = (-np.arange(nx)) % nx
    iy = (-np.arange(ny)) % ny
    iz = (-np.arange(nz)) % nz
    s_neg = np.empty_like(s)
    s_neg[np.ix_(ix, iy, iz)] = s
    o1r, o1i = J.green(s.real.reshape(-1, 3, 3), s.imag.reshape(-1, 3, 3))
    o2r, o2i = J.green(s_neg.real.reshape(-1, 3, 3), s_neg.imag.reshape(-1, 3, 3))
    out1 = (o1r + 1j * o1i).reshape(nx, ny, nz, 3, 3)
    out2 = (o2r + 1j * o2i).reshape(nx, ny, nz, 3, 3)
    splits = 0
    worst = {0: 0.0, 1: 0.0}
    for x in range(nx):
        for y in range(ny):
            for kz in range(nz // 2 + 1):
                if x + y + kz == 0:
                    continue
                s12 = np.array([[s[x, y, kz, a_, b_].real, s[x, y, kz, a_, b_].imag] for a_, b_ in COMP]).ravel()
                for fp32 in (0, 1):
                    e12, w6 = np.zeros(12), np.zeros(6)
                    split = hostlib.gpc_green_point(fp32, ga.ctypes.data_as(O.DP), x, y, kz, nx, ny, nz,
                                                    s12.ctypes.data_as(O.DP), e12.ctypes.data_as(O.DP),
                                                    w6.ctypes.data_as(O.DP))
                    ref = out1[x, y, kz]
                    if split:
                        ref = 0.5 * (ref + out2[ix[x], iy[y], iz[kz]])
                    sym = np.array([0.5 * (ref[i, j] + ref[j, i]) for i, j in COMP])
                    anti = np.array([0.5 * (ref[i, j] - ref[j, i]) for i, j in ANTI])
                    got_sym = e12[0::2] + 1j * e12[1::2]
                    got_anti = w6[:3] + 1j * w6[3:]
                    scale = np.max(np.abs(ref))
                    err = max(np.max(np.abs(got_sym - sym)), np.max(np.abs(got_anti - anti))) / scale
                    worst[fp32] = max(worst[fp32], err)
                    if not fp32:
                        splits += split
    assert worst[0] < 1e-13, worst
    assert worst[1] < 2e-6, worst
    # the even lengths put split entries on the half spectrum
    assert splits > 0
