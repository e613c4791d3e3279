"""Per-voxel pieces of the adjoint (m-themes_amd/csrc/adjoint_voxel.h, host build
in tests/hostcheck).

Each transposed linearisation is checked with the dot-product identity
<ybar, J dx> = <J^T ybar, dx>, where J dx comes from central finite
differences of the forward function (the same source file), on random
inputs of the sizes the loop uses (12 slip systems, Cu-like CRSS).  The
forward functions themselves are pinned against the GPU forward by the
loop-level gradient tests.
"""
import ctypes as C
import os

import numpy as np
import pytest

import subprocess

DP = C.POINTER(C.c_double)
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hostcheck")
LIB = os.path.join(HERE, "lib", "libhostcheck.so")


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    lib = C.CDLL(LIB)
    lib.adjc_voce.restype = C.c_double
    lib.adjc_voce.argtypes = [C.c_double] * 6
    lib.adjc_voce_grad.argtypes = [C.c_double] * 6 + [DP]
    lib.adjc_tangent.argtypes = [DP, DP, C.c_double, DP]
    lib.adjc_harden.argtypes = [C.c_int, C.c_double, DP, DP, DP, DP, DP, DP]
    lib.adjc_harden_bar.argtypes = [C.c_int, C.c_double, DP, DP, DP, DP, C.c_double, DP, DP, DP, DP]
    lib.adjc_orient.argtypes = [C.c_int, C.c_double, DP, DP, DP, DP, DP, DP]
    lib.adjc_orient_bar.argtypes = [C.c_int, C.c_double, DP, DP, DP, DP, DP, DP, DP, DP, DP]
    lib.adjc_rates.argtypes = [C.c_int, DP, DP, DP, DP, DP, DP, DP, DP]
    lib.adjc_rates_bar.argtypes = [C.c_int, DP, DP, DP, DP, DP, DP, DP, DP, DP, DP, DP]
    lib.adjc_schmid.argtypes = [DP, DP, DP, DP]
    lib.adjc_schmid_bar.argtypes = [DP, DP, DP, DP, DP]
    return lib


def p(a):
    return a.ctypes.data_as(DP)


def arr(*shape):
    return np.zeros(shape, dtype=np.float64)


def rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def fcc(rng):
    """The 12 {111}<110> systems, unit normals and directions."""
    n = []
    b = []
    for nn in [(1, 1, 1), (-1, 1, 1), (1, -1, 1), (1, 1, -1)]:
        nn = np.array(nn, float)
        for bb in [(1, -1, 0), (1, 0, -1), (0, 1, -1), (1, 1, 0), (1, 0, 1), (0, 1, 1)]:
            bb = np.array(bb, float)
            if abs(nn @ bb) < 1e-12:
                n.append(nn / np.sqrt(3))
                b.append(bb / np.sqrt(2))
    n = np.array(n)
    b = np.array(b)
    assert n.shape == (12, 3)
    assert np.allclose(np.sum(n * b, axis=1), 0)
    return np.ascontiguousarray(n), np.ascontiguousarray(b)


def check(lhs, rhs, tol):
    assert abs(lhs - rhs) <= tol * max(abs(lhs), abs(rhs), 1e-300), (lhs, rhs)


def test_schmid_bar(L):
    rng = np.random.default_rng(1)
    T = rotation(rng).ravel().copy()
    n, b = fcc(rng)
    cb = rng.normal(size=5)
    dT = rng.normal(size=9)
    h = 1e-6

    def f(t):
        c = arr(5)
        L.adjc_schmid(p(t), p(n[3].copy()), p(b[3].copy()), p(c))
        return cb @ c

    fd = (f(T + h * dT) - f(T - h * dT)) / (2 * h)
    Tb = arr(9)
    L.adjc_schmid_bar(p(T), p(n[3].copy()), p(b[3].copy()), p(cb), p(Tb))
    check(fd, Tb @ dT, 1e-8)


def _rate_inputs(rng):
    T = rotation(rng)
    n, b = fcc(rng)
    c = arr(12, 5)
    for s in range(12):
        cc = arr(5)
        P = 0.5 * (np.outer(T @ n[s], T @ b[s]) + np.outer(T @ b[s], T @ n[s]))
        r2, r6 = 1 / np.sqrt(2), 1 / np.sqrt(6)
        cc[:] = [-r2 * P[0, 0] + r2 * P[1, 1], -r6 * P[0, 0] - r6 * P[1, 1] + 2 * r6 * P[2, 2],
                 2 * r2 * P[1, 2], 2 * r2 * P[0, 2], 2 * r2 * P[0, 1]]
        c[s] = cc
    crss = 16.0 + 4.0 * rng.random(24)
    nrs = np.full(12, 10.0) + 0.3 * rng.random(12)
    g0 = np.ones(12)
    sig = rng.normal(size=6) * 25.0
    return c, crss, nrs, g0, sig


@pytest.mark.parametrize("wrt", ["sig", "c", "crss", "n"])
def test_rates_bar(L, wrt):
    rng = np.random.default_rng({"sig": 2, "c": 3, "crss": 4, "n": 5}[wrt])
    c, crss, nrs, g0, sig = _rate_inputs(rng)
    gdb = rng.normal(size=12)
    edb = np.append(rng.normal(size=5), 0.0)

    def f(c_, crss_, nrs_, sig_):
        gd, ed = arr(12), arr(6)
        L.adjc_rates(12, p(np.ascontiguousarray(c_)), p(crss_), p(nrs_), p(g0), p(sig_), p(gd), p(ed), None)
        return gdb @ gd + edb @ ed

    x = {"sig": sig, "c": c, "crss": crss, "n": nrs}[wrt]
    dx = rng.normal(size=x.shape)
    h = 1e-6 * max(1.0, np.abs(x).max())
    args = dict(c_=c, crss_=crss, nrs_=nrs, sig_=sig)
    key = {"sig": "sig_", "c": "c_", "crss": "crss_", "n": "nrs_"}[wrt]
    ap = dict(args); ap[key] = x + h * dx
    am = dict(args); am[key] = x - h * dx
    fd = (f(**ap) - f(**am)) / (2 * h)
    sigb, cb, crssb, nb = arr(6), arr(12, 5), arr(24), arr(12)
    L.adjc_rates_bar(12, p(c), p(crss), p(nrs), p(g0), p(sig), p(gdb), p(edb), p(sigb), p(cb), p(crssb), p(nb))
    xb = {"sig": sigb, "c": cb, "crss": crssb, "n": nb}[wrt]
    check(fd, float(np.sum(xb * dx)), 1e-6)


def test_rates_jacobian_matches_fd(L):
    """d(edot)/d(sigma) of the rate law (the 5x5 block the tangent uses)."""
    rng = np.random.default_rng(6)
    c, crss, nrs, g0, sig = _rate_inputs(rng)
    gd, ed, dd = arr(12), arr(6), arr(36)
    L.adjc_rates(12, p(c), p(crss), p(nrs), p(g0), p(sig), p(gd), p(ed), p(dd))
    J = np.zeros((6, 6))
    for j in range(6):
        h = 1e-6 * 25
        e1, e2 = arr(6), arr(6)
        s1, s2 = sig.copy(), sig.copy()
        s1[j] += h
        s2[j] -= h
        L.adjc_rates(12, p(c), p(crss), p(nrs), p(g0), p(s1), p(arr(12)), p(e1), None)
        L.adjc_rates(12, p(c), p(crss), p(nrs), p(g0), p(s2), p(arr(12)), p(e2), None)
        J[:, j] = (e1 - e2) / (2 * h)
    assert np.abs(J - dd.reshape(6, 6)).max() <= 1e-6 * np.abs(J).max()


@pytest.mark.parametrize("G,dg", [(0.0, 1e-3), (0.02, 3e-3), (0.5, 1e-4)])
def test_voce_grad(L, G, dg):
    x0 = np.array([G, dg, 10.0, 200.0, 20.0])  # G, dg, tau1, theta0, theta1
    d = arr(5)
    L.adjc_voce_grad(G, dg, 16.0, 10.0, 200.0, 20.0, p(d))
    for i in range(5):
        h = 1e-6 * max(abs(x0[i]), 1e-3)
        xp, xm = x0.copy(), x0.copy()
        xp[i] += h
        xm[i] -= h
        fp = L.adjc_voce(xp[0], xp[1], 16.0, xp[2], xp[3], xp[4])
        fm = L.adjc_voce(xm[0], xm[1], 16.0, xm[2], xm[3], xm[4])
        check((fp - fm) / (2 * h), d[i], 2e-6)


def _hard_inputs(rng):
    ns = 12
    hard = np.ones((ns, ns)) + 0.2 * rng.random((ns, ns))
    tau = np.tile([16.0, 16.0, 10.0], (ns, 1)) + 0.5 * rng.random((ns, 3))
    thet = np.tile([200.0, 20.0], (ns, 1)) + rng.random((ns, 2))
    gd = rng.normal(size=ns) * 0.3
    return ns, hard, tau, thet, gd


@pytest.mark.parametrize("wrt", ["gd", "G", "tau1", "theta0", "theta1"])
def test_harden_bar(L, wrt):
    rng = np.random.default_rng(7)
    ns, hard, tau, thet, gd = _hard_inputs(rng)
    dt = 1e-3
    G0 = 0.013
    crssb = rng.normal(size=24)
    Gb0 = rng.normal()

    def f(gd_, G_, tau_, thet_):
        crss = np.full(24, 20.0)
        G = C.c_double(G_)
        L.adjc_harden(ns, dt, p(hard), p(np.ascontiguousarray(tau_)), p(np.ascontiguousarray(thet_)), p(gd_),
                      p(crss), C.byref(G))
        return crssb @ crss + Gb0 * G.value

    Gb = C.c_double(Gb0)
    gdb, pb = arr(12), arr(3)
    L.adjc_harden_bar(ns, dt, p(hard), p(tau), p(thet), p(gd), G0, p(crssb), C.byref(Gb), p(gdb), p(pb))
    if wrt == "gd":
        dx = rng.normal(size=12)
        h = 1e-7
        fd = (f(gd + h * dx, G0, tau, thet) - f(gd - h * dx, G0, tau, thet)) / (2 * h)
        check(fd, gdb @ dx, 1e-6)
    elif wrt == "G":
        h = 1e-7
        fd = (f(gd, G0 + h, tau, thet) - f(gd, G0 - h, tau, thet)) / (2 * h)
        check(fd, Gb.value, 1e-6)
    else:
        # scalar parameter applied to every system: the sum of per-system partials
        i = {"tau1": 0, "theta0": 1, "theta1": 2}[wrt]
        tp, tm, hp, hm = tau.copy(), tau.copy(), thet.copy(), thet.copy()
        h = 1e-5
        if i == 0:
            tp[:, 2] += h
            tm[:, 2] -= h
        else:
            hp[:, i - 1] += h
            hm[:, i - 1] -= h
        fd = (f(gd, G0, tp, hp) - f(gd, G0, tm, hm)) / (2 * h)
        check(fd, pb[i], 1e-6)


@pytest.mark.parametrize("small", [False, True])
def test_orient_bar(L, small):
    rng = np.random.default_rng(8 + small)
    n, b = fcc(rng)
    ns = 12
    dt = 1e-3
    udot_a = np.zeros(9)
    udot_a[1], udot_a[3] = 0.1, -0.1
    T = rotation(rng).ravel().copy()
    gd = rng.normal(size=ns) * (1e-4 if small else 0.5)
    dW = rng.normal(size=3) * (1e-10 if small else 1e-3)
    if small:
        udot_a[:] = 0.0
    Tb0 = rng.normal(size=9)
    dT, dgd, ddW = rng.normal(size=9), rng.normal(size=ns), rng.normal(size=3)

    def f(T_, dW_, gd_):
        t = T_.copy()
        L.adjc_orient(ns, dt, p(udot_a), p(n), p(b), p(t), p(dW_), p(gd_))
        return Tb0 @ t

    Tb, gdb, dWb = Tb0.copy(), arr(12), arr(3)
    L.adjc_orient_bar(ns, dt, p(udot_a), p(n), p(b), p(T), p(dW), p(gd), p(Tb), p(gdb), p(dWb))
    h = 1e-7
    if small:
        ddW *= 1e-3  # |omega| stays below 1e-6 (the small-angle branch) at +-h
    fd = (f(T + h * dT, dW + h * ddW, gd + h * dgd) - f(T - h * dT, dW - h * ddW, gd - h * dgd)) / (2 * h)
    check(fd, Tb @ dT + dWb @ ddW + gdb @ dgd, 1e-6)


def test_tangent_is_inverse(L):
    rng = np.random.default_rng(9)
    A = rng.normal(size=(6, 6))
    S = A @ A.T + 6 * np.eye(6)
    B = rng.normal(size=(5, 5))
    D = np.zeros((6, 6))
    D[:5, :5] = B @ B.T
    Ct = arr(36)
    assert L.adjc_tangent(p(S.ravel().copy()), p(D.ravel().copy()), 1e-3, p(Ct)) == 0
    ref = np.linalg.inv(S + 1e-3 * D)
    assert np.abs(Ct.reshape(6, 6) - ref).max() <= 1e-12 * np.abs(ref).max()
    # not positive definite -> refused
    assert L.adjc_tangent(p((-S).ravel().copy()), p(D.ravel().copy()), 1e-3, p(Ct)) == 1


def test_adjoint_fixture_is_consistent():
    """The oracle gradient fixture: loss and gradient agree with its own curve,
    target and Jacobian (dL/dp = 2 sum_k (S_k - t_k) dS_k/dp)."""
    from conftest import GOLDEN
    fx = np.load(os.path.join(GOLDEN, "oracle_adjoint_8.npz"))
    r = fx["curve"][:, 2] - fx["target"]
    assert abs(np.sum(r ** 2) - fx["loss"]) <= 1e-12 * fx["loss"]
    g = 2.0 * np.einsum("k,ki->i", r, fx["jac_fd"][:, 2, :])
    assert np.all(np.abs(g - fx["grad_fd"]) <= 1e-6 * np.abs(fx["grad_fd"]))
    # hardening parameters act from the fourth increment only (harden after istep > 1)
    assert np.all(fx["jac_fd"][:3, 2, 1:4] == 0.0)
