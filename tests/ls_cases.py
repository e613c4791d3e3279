"""Single-voxel inputs of the per-voxel stress update that drive its line
search (src/evolution.c:166-253) through every branch natural inputs reach:
accepted full steps, TOLX rejections, quadratic backtracks and cubic
backtracks with b <= 0 and b > 0.  Shared by the oracle pins
(test_oracle_line_search.py) and the GPU parity test (test_gpu_parity.py).

Random orientation, trial stress lambda and strain over several decades,
zero plastic strain, equal forward/backward CRSS; the rate exponent and the
CRSS take a few values (steep rate laws backtrack).
"""
import numpy as np
from scipy.spatial.transform import Rotation

from pyevp import synth

NRS_VALUES = (10.0, 20.0, 30.0)


def cubic_t4(c11, c12, c44):
    c = np.zeros((3, 3, 3, 3))
    for i in range(3):
        c[i, i, i, i] = c11
        for j in range(3):
            if i != j:
                c[i, i, j, j] = c12
                c[i, j, i, j] = c[i, j, j, i] = c44
    return c


def slip_vectors():
    nb = np.array([s[0] for s in synth.FCC_SLIP], float)
    bb = np.array([s[1] for s in synth.FCC_SLIP], float)
    return nb / np.linalg.norm(nb, axis=1)[:, None], bb / np.linalg.norm(bb, axis=1)[:, None]


def cases(n, seed, nrs):
    """n voxels for rate exponent nrs: (T, lam, strain, crss) per voxel;
    T crystal -> sample, lam and strain symmetric 3x3, crss (12, 2)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        T = Rotation.random(random_state=int(rng.integers(1 << 30))).as_matrix()
        lam = rng.standard_normal((3, 3)) * rng.choice([1e-2, 0.1, 1, 10, 100, 1000])
        lam = lam + lam.T
        E = rng.standard_normal((3, 3)) * rng.choice([1e-6, 1e-5, 1e-4, 1e-3, 1e-2, 3e-2])
        E = E + E.T
        crss = np.full((12, 2), rng.choice([0.1, 1.0, 4.0, 16.0]))
        out.append((T, lam, E, crss))
    return out
