"""FP64 reference gradient for the adjoint tests, from the CPU oracle.

Job: 8^3 periodic voxels, 8 Voronoi grains (seed 0x5EED0008), the Cu
material and uniaxial-tension loading of the synthetic configurations,
5 increments of dt = 1e-3 (orientation update and hardening act from the
third increment on, so hardening reaches the stress from the fourth).

Validation tolerances: Newton step 1e-14 and line-search TOLX 1e-13 (the
reference's 1e-10 / 1e-7 reject Newton steps below 1e-7 relative, which
quantises the loop output at about that level, too coarse for a 1e-6
finite-difference step), Green tolerance 1e-11.  The converged equations
are the same; only how tightly they are solved changes.

Loss: L(p) = sum_k (Sigma_33,k(p) - target_k)^2, with the target curve from
a run at p * (1 + delta), delta = +5 % on each of p = (tau0, tau1, theta0,
theta1, n).  Gradient: central differences with step 1e-6 |p_i|, each
perturbed run forced to the base run's per-increment iteration counts (the
loop output is then a smooth function of p; a count that flips between the
+h and -h runs would add a jump of the size of the convergence tolerance).

    python tests/golden/make_adjoint_golden.py
"""
import dataclasses
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "m-themes_amd"))

import oracle as O  # noqa: E402
from pyevp import synth  # noqa: E402

P0 = np.array([16.0, 10.0, 200.0, 20.0, 10.0])
DELTA = 0.05
NINC = 5
H_REL = 1e-6


NEWTON_TOL, TOLX, ERR = 1e-14, 1e-13, 1e-11


def config(p):
    m = synth.Material(tau0f=p[0], tau0b=p[0], tau1=p[1], thet0=p[2], thet1=p[3], nrs=p[4])
    return synth.Config(n=8, grains=8, seed=0x5EED0008, increments=NINC, err=ERR, iter_max=40000,
                        material=m)


def curve(p, d, forced=None):
    """Sigma (n x 6: 11 22 33 23 13 12) after each increment, and the iteration counts."""
    inp = synth.write_job(d, config(p))
    J = O.Job(inp)
    out, iters = [], []
    for k in range(NINC):
        f = None if forced is None else np.array([forced[k]], dtype=np.int32)
        iters.append(int(J.run(1, f)[0]))
        s = J.macro()["SigAvg"].reshape(3, 3)
        out.append([s[0, 0], s[1, 1], s[2, 2], s[1, 2], s[0, 2], s[0, 1]])
    return np.array(out), np.array(iters)


def main(path=os.path.join(HERE, "oracle_adjoint_8.npz")):
    O.set_newton_tol(NEWTON_TOL, TOLX)
    with tempfile.TemporaryDirectory() as d:
        target, it_t = curve(P0 * (1 + DELTA), os.path.join(d, "t"))
        base, iters = curve(P0, os.path.join(d, "b"))
        loss = float(np.sum((base[:, 2] - target[:, 2]) ** 2))
        grad = np.zeros(5)
        jac = np.zeros((NINC, 6, 5))  # d Sigma_k,c / d p_i
        for i in range(5):
            h = H_REL * abs(P0[i])
            pp, pm = P0.copy(), P0.copy()
            pp[i] += h
            pm[i] -= h
            cp, _ = curve(pp, os.path.join(d, f"p{i}"), iters)
            cm, _ = curve(pm, os.path.join(d, f"m{i}"), iters)
            lp = np.sum((cp[:, 2] - target[:, 2]) ** 2)
            lm = np.sum((cm[:, 2] - target[:, 2]) ** 2)
            grad[i] = (lp - lm) / (2 * h)
            jac[:, :, i] = (cp - cm) / (2 * h)
            print(f"param {i}: dL/dp = {grad[i]:.10e}", flush=True)
    np.savez_compressed(path, params=P0, delta=DELTA, target=target[:, 2], curve=base, iters=iters,
                        target_iters=it_t, loss=loss, grad_fd=grad, jac_fd=jac, h_rel=H_REL,
                        newton_tol=NEWTON_TOL, tolx=TOLX, err=ERR)
    print(path, "iters", iters, "loss", loss, flush=True)


if __name__ == "__main__":
    main()
