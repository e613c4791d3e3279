"""Pins of the oracle's Newton iterate path: the NR_MODIFIED line search of
UpdateStress (src/evolution.c:166-253) and the strain-rate lag (Q5,
src/evolution.c:116-125,224-225,292-296).

The oracle records every trial point of a single-voxel solve (orc_ls_trace).
Each record is checked against an independent numpy restatement of the
residual R(sigma) = sigma - lambda + C0 (eps_p + dt edot(sigma) + S sigma - E)
(StrainRate_eval written from its closed form, the Schmid tensors built from
the slip vectors and the rotation):

* the merit values f = |R|^2 / 2 at 0 and at every trial point;
* the step is the Newton direction: J step = -R with J the Jacobian of the
  numpy R from the power law's closed-form derivative, itself checked against
  central differences (scaled to |step| = stpmax = 10 when capped);
* the slope is the derivative of lambda -> f(sigma + lambda step) at 0;
* a trial is accepted exactly when it meets the sufficient-decrease
  condition f <= f0 + 1e-4 lambda slope, rejected (step undone) exactly when
  lambda < TOLX / max_i |step_i| / max(|sigma_i|, 1);
* each backtrack is the minimiser of the model through the known values:
  the quadratic through f0, slope, f(1) on the first, the cubic through
  f0, slope and the last two trials afterwards (coefficients from a linear
  solve, stationary points from numpy.roots) capped at 0.5 of the last step,
  and at least 0.1 of it;
* the Newton loop stops at the first iteration whose accepted change is at
  most 1e-10 |lambda|;
* edot is the rate at the stress the last Newton iteration started from, and
  gamdot the rate at its last trial point (Q5).

Two cubic branches are not reached by these voxels: |a| < 1e-4 (rare) and
the cubic without a stationary point (unreachable after a rejected trial,
shown in test_backtracking_rule_on_every_branch).  The backtracking rule is
checked on every branch with synthetic values, for the oracle against the
models and for the Newton kernel's own rule (m-themes_amd/csrc/linesearch.h,
host build) against the oracle's.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import ls_cases

ROW = 24
HERE = os.path.dirname(os.path.abspath(__file__))
HOSTCHECK = os.path.join(HERE, "hostcheck", "lib", "libhostcheck.so")
DT = 1e-3
STPMAX, ALF, TOLX = 10.0, 1e-4, 1e-7


def _lib():
    L = O.lib()
    L.orc_ls_trace.argtypes = [O.DP, C.c_int]
    L.orc_ls_trace_count.restype = C.c_int
    L.orc_ls_next_lambda.restype = C.c_double
    L.orc_ls_next_lambda.argtypes = [C.c_double] * 6 + [O.IP]
    return L


class Voxel:
    """The numpy restatement of one voxel's residual."""

    def __init__(self, T, lam, E, crss, nrs, c066):
        self.B = O.basis()
        nb, bb = ls_cases.slip_vectors()
        n_s = nb @ T.T
        b_s = bb @ T.T
        self.P = 0.5 * (np.einsum("si,sj->sij", b_s, n_s) + np.einsum("si,sj->sij", n_s, b_s))
        C4 = ls_cases.cubic_t4(168400.0, 121400.0, 75400.0)
        Cgr = np.einsum("ip,jq,kr,ls,pqrs->ijkl", T, T, T, T, C4)
        self.Cgr66 = np.einsum("iab,abcd,jcd->ij", self.B, Cgr, self.B)
        self.S = np.linalg.inv(self.Cgr66)
        self.C0 = c066
        self.lam6 = self.to6(lam)
        self.E6 = self.to6(E)
        self.crss = crss
        self.nrs = nrs
        self.T = T

    def to6(self, t):
        return np.einsum("iab,ab->i", self.B, t)

    def from6(self, v):
        return np.einsum("i,iab->ab", v, self.B)

    def rates(self, s6):
        sig = self.from6(s6)
        tau = np.einsum("sij,ij->s", self.P, sig)
        tc = np.where(tau < 0, self.crss[:, 1], self.crss[:, 0])
        x = tau / tc
        gd = np.abs(x) ** self.nrs * np.sign(x)
        return self.to6(np.einsum("sij,s->ij", self.P, gd)), gd

    def R(self, s6):
        ed, _ = self.rates(s6)
        return s6 - self.lam6 + self.C0 @ (ed * DT + self.S @ s6 - self.E6)

    def f(self, s6):
        r = self.R(s6)
        return 0.5 * r @ r

    def rounding(self, s6):
        """Bound on the rounding of R at s6: a few hundred ulps of each term it
        sums; the rate term's relative rounding is the resolved shear
        stresses' (|sigma| / |tau| ulps, from cancellation) times n."""
        ed, _ = self.rates(s6)
        tau = np.abs(np.einsum("sij,ij->s", self.P, self.from6(s6))).max()
        cond = max(1.0, np.abs(s6).max() / max(tau, 1e-300))
        lin = [s6, self.lam6, self.C0 @ (self.S @ s6), self.C0 @ self.E6]
        u = 256 * np.finfo(float).eps
        return u * (sum(np.abs(t).max() for t in lin) +
                    np.abs(self.C0 @ ed * DT).max() * max(self.nrs, 1.0) * cond)

    def rate_rounding(self, s6):
        """Relative rounding bound of the rates at s6 (|sigma| / |tau| ulps of
        the resolved shear stresses, times n)."""
        tau = np.abs(np.einsum("sij,ij->s", self.P, self.from6(s6))).max()
        cond = max(1.0, np.abs(s6).max() / max(tau, 1e-300))
        return 256 * np.finfo(float).eps * max(self.nrs, 1.0) * cond

    def jac(self, s6):
        """dR/dsigma from the closed-form derivative of the power law:
        I + C0 (S + dt sum_s P_s (x) P_s n |x|^(n-1) / tau_c)."""
        sig = self.from6(s6)
        tau = np.einsum("sij,ij->s", self.P, sig)
        tc = np.where(tau < 0, self.crss[:, 1], self.crss[:, 0])
        dg = self.nrs * np.abs(tau / tc) ** (self.nrs - 1) / tc
        P6 = np.einsum("iab,sab->si", self.B, self.P)
        D = np.einsum("si,sj,s->ij", P6, P6, dg)
        return np.eye(6) + self.C0 @ (self.S + DT * D)

    def jac_fd(self, s6):
        J = np.zeros((6, 6))
        for j in range(6):
            h = 1e-6 * max(abs(s6[j]), 1.0)
            e = np.zeros(6)
            e[j] = h
            J[:, j] = (self.R(s6 + e) - self.R(s6 - e)) / (2 * h)
        return J


def solve(vx, L):
    """Oracle solve of one voxel with the trace on: (status, sig, edot, gamdot, n_iter, rows)."""
    nb, bb = ls_cases.slip_vectors()
    xt = np.array([O.t2_to_v5(0.5 * (np.outer(nb[s], bb[s]) + np.outer(bb[s], nb[s])))
                   for s in range(12)])
    sc = O.rotate_schmid(vx.T, xt)
    buf = np.zeros((4096, ROW))
    L.orc_ls_trace(O.ptr(buf), 4096)
    try:
        st, s, ed, gd, _, _, nit = O.update_stress_voxel(
            O.lu_inv66(vx.Cgr66), vx.C0, DT, vx.from6(vx.lam6), np.zeros((3, 3)), vx.from6(vx.E6),
            sc, vx.crss, np.full(12, vx.nrs), np.ones(12))
    finally:
        n = L.orc_ls_trace_count()
        L.orc_ls_trace(None, 0)
    return st, s, ed, gd, nit, buf[:n].copy()


def model_minimiser(lam1, f1, lam2, f2, f0, slope):
    """Minimiser of the model a line search fits after a rejected trial at
    lam1 (value f1); lam2/f2 the trial before it (None on the first)."""
    if lam2 is None:
        c = f1 - f0 - slope * lam1
        r = np.roots([2 * c, slope])
        return float(r[0].real), "quadratic"
    A = np.array([[lam1 ** 3, lam1 ** 2], [lam2 ** 3, lam2 ** 2]])
    a, b = np.linalg.solve(A, [f1 - f0 - slope * lam1, f2 - f0 - slope * lam2])
    if abs(a) < 1e-4:
        return float(-slope / (2 * b)), "quadratic part"
    r = np.roots([3 * a, 2 * b, slope])
    real = [x.real for x in r if abs(x.imag) <= 1e-12 * max(1.0, abs(x.real))]
    if not real:
        return 0.5 * lam1, "no stationary point"
    mins = [x for x in real if 6 * a * x + 2 * b > 0]
    assert len(mins) == 1
    return float(mins[0]), "cubic"


def collect(nrs, n, seed, c066):
    L = _lib()
    out = []
    for (T, lam, E, crss) in ls_cases.cases(n, seed, nrs):
        vx = Voxel(T, lam, E, crss, nrs, c066)
        st, s, ed, gd, nit, rows = solve(vx, L)
        if st:
            continue
        out.append((vx, s, ed, gd, nit, rows))
    return out


@pytest.fixture(scope="module")
def solved():
    c066 = O.t4_to_c2(ls_cases.cubic_t4(168400.0, 121400.0, 75400.0))
    res = []
    for k, nrs in enumerate(ls_cases.NRS_VALUES):
        res += collect(nrs, 120, 100 + k, c066)
    return res


def test_cases_reach_every_natural_branch(solved):
    oc = np.concatenate([r[5][:, 6] for r in solved]).astype(int)
    counts = np.bincount(oc, minlength=7)
    # accepted, TOLX-rejected, quadratic, cubic with b <= 0 and with b > 0
    for k in (0, 1, 2, 5, 6):
        assert counts[k] > 0, counts


def test_merit_values_step_and_slope(solved):
    checked = 0
    for vx, s, ed, gd, nit, rows in solved:
        seen = set()
        for r in rows:
            it = int(r[0])
            old, step = r[8:14], r[14:20]
            # |R| = sqrt(2 f), compared within the rounding of R's terms
            for lam, fo in ((0.0, r[3]), (r[1], r[2])):
                pt = old + lam * step
                rn = np.linalg.norm(vx.R(pt))
                assert abs(np.sqrt(2 * fo) - rn) <= vx.rounding(pt), (fo, rn)
            if it in seen:
                continue
            seen.add(it)
            Rv = vx.R(old)
            dR = vx.rounding(old)
            if np.linalg.norm(Rv) < 1e3 * dR:
                continue  # a residual at rounding level fixes no direction
            # Newton direction: J step = -R (step scaled to stpmax when capped)
            full = step * (r[21] / STPMAX) if r[22] else step
            J = vx.jac(old)
            if np.linalg.cond(J) > 1e8:
                continue  # far in the power law's tail: the direction is rounding
            # the closed-form Jacobian against central differences of R
            Jf = vx.jac_fd(old)
            assert np.abs(J - Jf).max() <= 1e-5 * np.abs(J).max()
            # the step comes from an explicit LU inverse: its residual is
            # eps * cond(J) relative, not backward-stable
            scale = np.abs(J).max() * np.abs(full).max()
            tol = 256 * np.finfo(float).eps * np.linalg.cond(J) * scale + 10 * dR
            assert np.linalg.norm(J @ full + Rv) <= tol
            assert (r[22] == 1) == (np.linalg.norm(full) > STPMAX)
            # slope = d/dlam f(old + lam step) at 0
            h = 1e-5
            d = (vx.f(old + h * step) - vx.f(old - h * step)) / (2 * h)
            assert abs(r[4] - d) <= 1e-5 * abs(r[4]) + np.linalg.norm(Rv) * dR / h, (r[4], d)
            checked += 1
    assert checked > 100


def test_acceptance_rejection_and_backtracking(solved):
    n_bt = 0
    for vx, s, ed, gd, nit, rows in solved:
        prev = None
        for k, r in enumerate(rows):
            it, lam, fnew, f0, slope, alamin, oc, nxt = int(r[0]), r[1], r[2], r[3], r[4], r[5], int(r[6]), r[7]
            old, step = r[8:14], r[14:20]
            first = prev is None or int(prev[0]) != it
            if first:
                assert lam == 1.0
            test = np.max(np.abs(step) / np.maximum(np.abs(old), 1.0))
            assert np.isclose(alamin, TOLX / test, rtol=1e-14)
            if lam < alamin:
                assert oc == 1
            elif fnew <= f0 + ALF * lam * slope:
                assert oc == 0
            else:
                assert oc in (2, 3, 4, 5, 6)
                m, kind = model_minimiser(lam, fnew, None if first else prev[1], None if first else prev[2],
                                          f0, slope)
                # the 0.5 cap applies to the cubic steps; the quadratic one is
                # below 0.5 / (1 - ALF) by the rejection itself
                expect = m if first else min(m, 0.5 * lam)
                assert abs(nxt - expect) <= 1e-8 * lam, (oc, kind, nxt, expect)
                assert (oc == 2) == first
                # the next trial of the same iteration
                nr = rows[k + 1]
                assert int(nr[0]) == it
                assert nr[1] == max(nxt, 0.1 * lam)
                assert 0.1 * lam <= nr[1] <= 0.5 * lam / (1 - ALF)
                n_bt += 1
            prev = r
    assert n_bt > 10


def test_newton_stop_rule_final_stress_and_rate_lag(solved):
    for vx, s, ed, gd, nit, rows in solved:
        its = rows[:, 0].astype(int)
        assert its[-1] == nit and np.all(np.diff(its) >= 0)
        lam_norm = np.linalg.norm(vx.from6(vx.lam6))
        starts = [rows[its == i][0, 8:14] for i in range(1, nit + 1)]
        last = rows[-1]
        oc = int(last[6])
        assert oc in (0, 1)
        end = last[8:14] + (last[1] * last[14:20] if oc == 0 else 0.0)
        ends = starts[1:] + [end]
        for i, (a, b) in enumerate(zip(starts, ends)):
            small = np.linalg.norm(b - a) / lam_norm <= 1e-10
            assert small == (i == nit - 1) or (nit == 500 and i == nit - 1)
        assert np.allclose(vx.to6(s), end, rtol=0, atol=1e-12 * max(np.abs(end).max(), 1e-300))
        # Q5: edot from the start of the last iteration, gamdot from its last
        # trial (the next-to-last iterate differs from the last by > 1e-10
        # relative, far above these bars)
        e_ref, _ = vx.rates(starts[-1])
        tol = vx.rate_rounding(starts[-1])
        assert np.abs(vx.to6(ed) - e_ref).max() <= tol * np.abs(e_ref).max() + 1e-300
        lt = last[8:14] + last[1] * last[14:20]
        _, g_ref = vx.rates(lt)
        tol = vx.rate_rounding(lt)
        assert np.abs(gd - g_ref).max() <= tol * np.abs(g_ref).max() + 1e-300


def _hostcheck():
    if not os.path.exists(HOSTCHECK):
        subprocess.run(["make", "-s", "-C", os.path.dirname(os.path.dirname(HOSTCHECK))], check=True)
    H = C.CDLL(HOSTCHECK)
    H.lsc_backtrack.restype = C.c_double
    H.lsc_backtrack.argtypes = [C.c_double] * 6
    return H


def test_backtracking_rule_on_every_branch():
    """The rule itself on synthetic values: the oracle against the model
    minimisers, the Newton kernel's rule (host build of linesearch.h)
    against the oracle's, bit for bit.

    After a rejected trial at lam, f(lam) > f0 + ALF lam slope, so the
    cubic's coefficients satisfy b lam + a lam^2 > (1 - ALF)|slope|; with
    a = -k |slope| / lam^2 < 0 that gives b^2 > (1 - ALF + k)^2 slope^2 / lam^2,
    and disc = b^2 - 3 a slope < 0 would need (1 - ALF + k)^2 < 3k, i.e.
    k^2 - (1 + 2 ALF) k + (1 - ALF)^2 < 0, which has no real solution for
    ALF < 1/4: the no-stationary-point branch is unreachable from the line
    search.  Values rejected trials can take reach the other four; the
    fifth is exercised with values a line search never passes."""
    L = _lib()
    H = _hostcheck()
    rng = np.random.default_rng(5)
    seen = np.zeros(7, int)
    oc = C.c_int(0)
    for _ in range(20000):
        f0 = 10.0 ** rng.uniform(-9, 3)
        slope = -2.0 * f0 * 10.0 ** rng.uniform(-3, 0)
        first = rng.random() < 0.2
        lam2 = 1.0 if first else 10.0 ** rng.uniform(-3, 0)
        lam = 1.0 if first else lam2 * rng.uniform(0.1, 0.5)
        # rejected trials: above the sufficient-decrease line
        f1 = f0 + ALF * lam * slope + abs(slope) * lam * 10.0 ** rng.uniform(-6, 2)
        f2 = f0 + ALF * lam2 * slope + abs(slope) * lam2 * 10.0 ** rng.uniform(-6, 2)
        t = L.orc_ls_next_lambda(lam, lam2, f1, f2, f0, slope, C.byref(oc))
        seen[oc.value] += 1
        m, kind = model_minimiser(lam, f1, None if first else lam2, None if first else f2, f0, slope)
        assert abs(t - (m if first else min(m, 0.5 * lam))) <= 1e-7 * lam, (oc.value, kind, t, m)
        assert {2: "quadratic", 3: "quadratic part", 5: "cubic", 6: "cubic"}[oc.value] == kind
        g = H.lsc_backtrack(lam, lam2, f1, f2, f0, slope)
        assert g == max(t, 0.1 * lam)
    assert np.all(seen[[2, 3, 5, 6]] > 0) and seen[4] == 0, seen
    # a cubic with no stationary point: a < 0, b^2 < 3 a slope (f(lam) below the line)
    for _ in range(200):
        f0 = 10.0 ** rng.uniform(-3, 3)
        slope = -f0 * 10.0 ** rng.uniform(-2, 0)
        lam2 = 10.0 ** rng.uniform(-2, 0)
        lam = lam2 * rng.uniform(0.1, 0.5)
        a = -10.0 ** rng.uniform(0, 3) * abs(slope)
        b = rng.uniform(-1, 1) * np.sqrt(3 * a * slope)
        f1 = f0 + slope * lam + b * lam ** 2 + a * lam ** 3
        f2 = f0 + slope * lam2 + b * lam2 ** 2 + a * lam2 ** 3
        t = L.orc_ls_next_lambda(lam, lam2, f1, f2, f0, slope, C.byref(oc))
        if abs(a) < 1e-4:
            continue
        assert oc.value == 4
        m, kind = model_minimiser(lam, f1, lam2, f2, f0, slope)
        assert kind == "no stationary point" and t == 0.5 * lam
        assert H.lsc_backtrack(lam, lam2, f1, f2, f0, slope) == max(t, 0.1 * lam)
