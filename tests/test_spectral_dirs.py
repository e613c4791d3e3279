"""The FP32 Green pass's wave-vector directions and Nyquist split test
(m-themes_amd/csrc/spectral.h, host build in tests/hostcheck) against the
FP64 ones, which follow the reference's unit-vector construction
(src/init.c:775-796) and the symmetrised operator's two directions at Nyquist
entries.

The FP32 form decides the split from the indices alone; it has to give the
FP64 comparison's answer at every wave vector, including mixed-radix and odd
lengths where k/N is inexact.  Its directions have to agree to FP32 precision
(the operator is even in n, so each is compared up to sign)."""
import ctypes as C
import itertools
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hostcheck")
LIB = os.path.join(HERE, "lib", "libhostcheck.so")


@pytest.fixture(scope="module")
def sweep():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    H = C.CDLL(LIB)
    H.ksc_sweep.restype = None
    H.ksc_sweep.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]

    def run(nx, ny, nz):
        out = (C.c_double * 4)()
        H.ksc_sweep(nx, ny, nz, out)
        return list(out)

    return run


SMALL = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 15, 16]
LARGE = [(128, 128, 128), (120, 96, 100), (48, 24, 24), (64, 50, 36), (30, 18, 20), (243, 6, 10)]


def test_split_rule_and_directions(sweep):
    tot = splits = 0
    worst = 0.0
    for dims in itertools.chain(itertools.product(SMALL, repeat=3), LARGE):
        n, s, bad, dmax = sweep(*dims)
        assert bad == 0, f"split decisions differ on {dims}"
        tot += n
        splits += s
        worst = max(worst, dmax)
    # the split case is exercised, and is a minority of wave vectors
    assert 0 < splits < tot
    # correctly rounded FP32 arithmetic on the host: a few ulp of 1
    assert worst <= 2e-7, worst


def test_split_cases_by_hand(sweep):
    # 2x2x2: every nonzero index is Nyquist, so n(-k) = n(k) everywhere
    assert sweep(2, 2, 2)[1] == 0
    # 4x1x1: k = 2 is Nyquist alone on its axis: no split; 4x3x1: (2, 1)
    # and (2, 2) mix a Nyquist and a non-Nyquist component: split
    n, s, bad, _ = sweep(4, 1, 1)
    assert (n, s, bad) == (3, 0, 0)
    n, s, bad, _ = sweep(4, 3, 1)
    assert (n, s, bad) == (11, 2, 0)
    # odd lengths have no Nyquist index
    assert sweep(5, 7, 9)[1] == 0
    assert np.isfinite(sweep(3, 3, 3)[3])
