"""The Green loop control's get_smacro step (m-themes_amd/csrc/loopctl_step.h,
the arithmetic the one-workgroup control kernel runs each iteration; host
build in tests/hostcheck) against the oracle's get_smacro core
(oracle/evp_oracle.c, src/evolution.c:473-547), bit for bit.

Both get the same column sums of the stress, so any difference is in the
step itself: the scaling to the averages, the mixed boundary condition's
macro strain correction dD (its lower entries kept), its running sum, and
the von Mises norms of the deviatoric averages."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hostcheck")
LIB = os.path.join(HERE, "lib", "libhostcheck.so")
COMP = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]


@pytest.fixture(scope="module")
def libs():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    H = C.CDLL(LIB)
    H.lcc_smacro.restype = None
    H.lcc_smacro.argtypes = [O.DP, O.DP, C.c_double, C.c_double, O.IP, O.IP, O.DP, O.DP, O.DP, O.DP]
    L = O.lib()
    L.orc_smacro_core.restype = None
    L.orc_smacro_core.argtypes = [O.DP, O.DP, C.c_double, C.c_double, O.DP, O.DP, O.IP, O.IP,
                                  O.DP, O.DP, O.DP, O.DP, O.DP, O.DP]
    return H, L


def _p(a, t=O.DP):
    return a.ctypes.data_as(t)


def _case(rng, vbc):
    n = int(rng.integers(8, 300)) ** 3
    c = dict(
        s0=rng.standard_normal(81) * 1e-5,
        scauchy=rng.standard_normal(9) * 50.0,
        wgt_ph1=float(rng.uniform(0.05, 1.0)),
        scale=1.0 / n,
        vbc=np.asarray(vbc, np.int32),
        sbc=(1 - np.asarray(vbc)).astype(np.int32),
        sums=rng.standard_normal(6) * 100.0 * n,
        sums1=rng.standard_normal(6) * 60.0 * n,
        dD=rng.standard_normal(9) * 1e-4,
        acum=rng.standard_normal(9) * 1e-3,
    )
    return c


def _device_step(H, c):
    red = np.zeros(16)
    red[3:9] = c["sums"]
    red[9:15] = c["sums1"]
    dD, acum, out = c["dD"].copy(), c["acum"].copy(), np.zeros(20)
    H.lcc_smacro(_p(c["s0"]), _p(c["scauchy"]), c["wgt_ph1"], c["scale"], _p(c["vbc"], O.IP),
                 _p(c["sbc"], O.IP), _p(red), _p(dD), _p(acum), _p(out))
    return dict(sig=out[:9], sig1=out[9:18], dD=dD, acum=acum, vm=out[18:20])


def _oracle_step(L, c):
    def full(s6):
        t = np.zeros((3, 3))
        for k, (a, b) in enumerate(COMP):
            t[a, b] = t[b, a] = s6[k]
        return t

    sum_, sum1 = full(c["sums"]), full(c["sums1"])
    sig, sig1 = np.zeros((3, 3)), np.zeros((3, 3))
    dD, acum = c["dD"].reshape(3, 3).copy(), c["acum"].reshape(3, 3).copy()
    vm, vm1 = C.c_double(), C.c_double()
    L.orc_smacro_core(_p(sum_), _p(sum1), c["scale"], c["wgt_ph1"], _p(c["s0"]), _p(c["scauchy"]),
                      _p(c["vbc"], O.IP), _p(c["sbc"], O.IP), _p(sig), _p(sig1), _p(dD), _p(acum),
                      C.byref(vm), C.byref(vm1))
    return dict(sig=sig.ravel(), sig1=sig1.ravel(), dD=dD.ravel(), acum=acum.ravel(),
                vm=np.array([vm.value, vm1.value]))


BCS = [
    [1, 1, 1, 1, 1, 1],  # strain rate prescribed everywhere: dD = 0
    [1, 1, 0, 1, 1, 1],  # uniaxial along z, lateral stresses free
    [0, 0, 1, 1, 1, 1],
    [0, 0, 0, 0, 0, 0],
]


@pytest.mark.parametrize("vbc", BCS)
def test_step_equals_oracle_bit_for_bit(libs, vbc):
    H, L = libs
    rng = np.random.default_rng(sum(v << i for i, v in enumerate(vbc)))
    for _ in range(200):
        c = _case(rng, vbc)
        d, o = _device_step(H, c), _oracle_step(L, c)
        for k in d:
            assert np.array_equal(d[k].view(np.uint64), o[k].view(np.uint64)), (k, d[k], o[k])


def test_step_outputs_mean_what_they_say(libs):
    """Values the step must produce whatever its arithmetic: the averages are
    the sums times the weights, dD vanishes under prescribed strain rates and
    is zero below the diagonal, and s_vm of a uniaxial average is |sigma_zz|."""
    H, _ = libs
    rng = np.random.default_rng(11)
    c = _case(rng, [1, 1, 1, 1, 1, 1])
    c["dD"][:] = 0.0
    d = _device_step(H, c)
    assert np.all(d["dD"] == 0.0)
    assert np.array_equal(d["acum"], c["acum"])
    s = c["sums"] * c["scale"]
    assert d["sig"][0] == s[0] and d["sig"][8] == s[2] and d["sig"][1] == d["sig"][3] == s[5]
    c = _case(rng, [1, 1, 0, 1, 1, 1])
    c["dD"][:] = 0.0
    c["sums"][:] = [0, 0, 250.0 / c["scale"], 0, 0, 0]
    d = _device_step(H, c)
    assert np.count_nonzero(d["dD"]) == 1 and d["dD"][8] != 0.0
    assert abs(d["vm"][0] - 250.0) < 1e-12 * 250.0
