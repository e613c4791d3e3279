"""Independent numpy statement of the uniform and laminated crystal problems
the Green loop reduces to, for the oracle pins of the plastic whole-loop
fields (test_oracle_plastic.py).

Written from the mechanics, not from the oracle's code:
* cubic stiffness rotated with the Bunge rotation (scipy), degrees converted
  with the reference's truncated PI (Q8, src/evp.h:77);
* Schmid tensors sym((T b) x (T n)) from the unit slip vectors, rate law
  gamdot_s = g0 |tau_s / tau_c|^n sgn(tau_s), edot = sum_s P_s gamdot_s;
* at convergence every layer satisfies S sigma + eps_p + dt edot(sigma) = E
  (the Newton residual at its fixed point), the layers of a laminate along x
  carry the same traction sigma . e_x and strains that differ by
  sym(c x e_x) (a displacement-gradient jump c x e_x), the mean strain
  component 33 grows by dt per increment, 23/13/12 stay at zero and the
  mean stresses 11 and 22 vanish;
* after the loop: eps_p += dt edot; from the third increment on (istep > 1
  with istep counted from 0, src/evolution.c:609) the lattice spin skew(L) - skew(sum_s (T b_s) x (T n_s) gamdot_s) rotates T by
  its exponential (scipy.linalg.expm; spins below 1e-6 per increment leave
  it unchanged, the reference's small-angle rule), and the CRSS follow the Voce law
  tau(G) = tau0 + (tau1 + theta1 G)(1 - exp(-G theta0 / tau1)) over the
  accumulated slip G (hardening matrix of ones).
"""
import numpy as np
from scipy.linalg import expm
from scipy.spatial.transform import Rotation

from ls_cases import cubic_t4, slip_vectors

PI_REF = 3.14159265359


def rotation(euler_deg):
    """Crystal -> sample rotation of Bunge angles in degrees."""
    return Rotation.from_euler("ZXZ", np.asarray(euler_deg) * PI_REF / 180.0).as_matrix()


def voce(G, tau0, tau1, th0, th1):
    return tau0 + (tau1 + th1 * G) * (1.0 - np.exp(-G * abs(th0 / tau1)))


def sym(a):
    return 0.5 * (a + a.T)


def skew(a):
    return 0.5 * (a - a.T)


class Layer:
    def __init__(self, euler_deg, mat):
        self.T = rotation(euler_deg)
        C4 = cubic_t4(mat.c11, mat.c12, mat.c44)
        T = self.T
        self.C = np.einsum("ip,jq,kr,ls,pqrs->ijkl", T, T, T, T, C4)
        # compliance on symmetric tensors: invert the 6x6 in an orthonormal
        # (Mandel) basis
        self.Sm = np.linalg.inv(to_mandel4(self.C))
        self.crss = np.full(12, mat.tau0f)
        self.G = 0.0
        self.epsp = np.zeros((3, 3))
        self.grad = np.zeros((3, 3))
        self.mat = mat

    def schmid(self):
        nb, bb = slip_vectors()
        n = nb @ self.T.T
        b = bb @ self.T.T
        return n, b, 0.5 * (np.einsum("si,sj->sij", b, n) + np.einsum("si,sj->sij", n, b))

    def rates(self, sig):
        _, _, P = self.schmid()
        tau = np.einsum("sij,ij->s", P, sig)
        x = tau / self.crss
        gd = self.mat.gam0 * np.abs(x) ** self.mat.nrs * np.sign(x)
        return np.einsum("sij,s->ij", P, gd), gd

    def strain_of(self, sig, dt):
        ed, _ = self.rates(sig)
        return from_mandel2(self.Sm @ to_mandel2(sig)) + self.epsp + dt * ed


MANDEL = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
W = np.array([1, 1, 1, np.sqrt(2), np.sqrt(2), np.sqrt(2)])


def to_mandel2(t):
    return np.array([t[i, j] for i, j in MANDEL]) * W


def from_mandel2(v):
    t = np.zeros((3, 3))
    for k, (i, j) in enumerate(MANDEL):
        t[i, j] = t[j, i] = v[k] / W[k]
    return t


def to_mandel4(C):
    M = np.zeros((6, 6))
    for a, (i, j) in enumerate(MANDEL):
        for b, (k, l) in enumerate(MANDEL):
            M[a, b] = C[i, j, k, l] * W[a] * W[b]
    return M


class Laminate:
    """One layer (uniform crystal) or two equal layers stacked along x."""

    def __init__(self, eulers, mat, dt=1e-3):
        self.L = [Layer(e, mat) for e in eulers]
        self.dt = dt
        self.Ebar = np.zeros((3, 3))  # DisGradAvg_actual
        self.istep = 0  # the reference counts increments from 0 (src/evolution.c:609)

    def solve_layer(self, lay, E, sig0):
        """sigma with S sigma + eps_p + dt edot(sigma) = E: Newton on the
        Mandel 6-vector with the closed-form tangent, halving the step until
        |r| decreases."""
        dt = self.dt
        _, _, P = lay.schmid()
        Pm = np.array([to_mandel2(p) for p in P])
        target = to_mandel2(E - lay.epsp)
        s = to_mandel2(sig0)

        def resid(s):
            tau = Pm @ s
            x = tau / lay.crss
            gd = lay.mat.gam0 * np.abs(x) ** lay.mat.nrs * np.sign(x)
            return lay.Sm @ s + dt * (Pm.T @ gd) - target, x

        r, x = resid(s)
        for _ in range(200):
            nr = np.linalg.norm(r)
            if nr <= 1e-17:
                break
            dg = lay.mat.gam0 * lay.mat.nrs * np.abs(x) ** (lay.mat.nrs - 1) / lay.crss
            J = lay.Sm + dt * np.einsum("si,sj,s->ij", Pm, Pm, dg)
            ds = -np.linalg.solve(J, r)
            a = 1.0
            while True:
                r2, x2 = resid(s + a * ds)
                if np.linalg.norm(r2) < nr or a < 1e-12:
                    break
                a *= 0.5
            if a < 1e-12:
                break
            s, r, x = s + a * ds, r2, x2
            if np.linalg.norm(a * ds) <= 1e-15 * np.linalg.norm(s):
                break
        return from_mandel2(s)

    def increment(self):
        dt = self.dt
        nl = len(self.L)
        Ud = np.zeros((3, 3))
        Ud[2, 2] = 1.0
        e1 = np.array([1.0, 0.0, 0.0])
        prev = [getattr(l, "sig", np.zeros((3, 3))) for l in self.L]

        def grads(y):
            base = self.Ebar + Ud * dt + np.diag([y[0], y[1], 0.0])
            if nl == 1:
                return [base]
            c = y[2:5]
            return [base + 0.5 * np.outer(c, e1), base - 0.5 * np.outer(c, e1)]

        def stresses(y):
            g = grads(y)
            return [self.solve_layer(lay, sym(g[l]), prev[l]) for l, lay in enumerate(self.L)]

        def F(y):
            sigs = stresses(y)
            mean = sum(sigs) / nl
            r = [mean[0, 0], mean[1, 1]]
            if nl == 2:
                r += list((sigs[0] - sigs[1]) @ e1)
            return np.array(r)

        # Newton on the macroscopic / interface unknowns from the previous
        # increment's values, halving the step until |F| decreases
        y = self.y.copy() if hasattr(self, "y") else np.zeros(2 if nl == 1 else 5)
        f = F(y)
        for _ in range(50):
            if np.max(np.abs(f)) < 1e-11:
                break
            J = np.zeros((len(y), len(y)))
            for j in range(len(y)):
                h = 1e-9
                e = np.zeros(len(y))
                e[j] = h
                J[:, j] = (F(y + e) - F(y - e)) / (2 * h)
            dy = -np.linalg.solve(J, f)
            a = 1.0
            while a > 1e-6:
                f2 = F(y + a * dy)
                if np.linalg.norm(f2) < np.linalg.norm(f):
                    break
                a *= 0.5
            y, f = y + a * dy, f2
        self.y = y
        res = float(np.max(np.abs(F(y))))
        sigs = stresses(y)
        g = grads(y)
        dE = np.diag([y[0], y[1], 0.0])
        out = []
        for l, lay in enumerate(self.L):
            ed, gd = lay.rates(sigs[l])
            velgrad = (g[l] - lay.grad) / dt
            lay.grad = g[l]
            lay.epsp = lay.epsp + dt * ed
            if self.istep > 1:
                n, b, _ = lay.schmid()
                lp = np.einsum("si,sj,s->ij", b, n, gd)
                rot = (skew(velgrad) - skew(lp)) * dt
                w = np.array([rot[2, 1], rot[0, 2], rot[1, 0]])
                # Orientation (src/constitutive.c:150-186) rotates by |w| about
                # w; below |w| = 1e-6 its small-angle rule turns by about
                # |w|^2 / 2, which leaves T unchanged to 1e-12
                if np.linalg.norm(w) >= 1e-6:
                    lay.T = expm(rot) @ lay.T
                dG = np.sum(np.abs(gd)) * dt
                m = lay.mat
                lay.crss = lay.crss + (voce(lay.G + dG, m.tau0f, m.tau1, m.thet0, m.thet1) -
                                       voce(lay.G, m.tau0f, m.tau1, m.thet0, m.thet1))
                lay.G += dG
            lay.sig = sigs[l]
            out.append(sigs[l])
        self.Ebar = self.Ebar + Ud * dt + dE
        self.istep += 1
        return out, res
