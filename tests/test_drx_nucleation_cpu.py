"""CPU checks of the DRX nucleation pieces.

The oracle's NucleationCheckDRX (oracle/pf_oracle.c, the shipped branch of
src/integrated.c:161-494) against hand-worked cases: the grain-boundary
indicator with the reference's -y in place of -z (src/constitutive.c:816),
the strict test rho_tot/12 > kappa, consecutive IDs in sweep order, spheres
without periodic wrap painted by the last nucleus.  The product's MT19937
(the host stream behind evpgpu_drx_weibull) against the C++ standard's
known answers for std::mt19937 (the same generator and seeding: first output
3499211612 and 10000th 4123659995 for seed 5489) and against numpy's
independent MT19937; the Weibull draws against the inverse of their CDF."""
import math
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
sys.path.insert(0, os.path.join(HERE, "..", "m-themes_amd"))

import oracle as o  # noqa: E402
from pyevp import pf  # noqa: E402


def test_mt19937_known_answers():
    raw = (pf.uniform_pos(5489, 10000) * 4294967296.0).astype(np.uint64)
    assert raw[0] == 3499211612 and raw[9999] == 4123659995


@pytest.mark.parametrize("seed", [1, 4357, 123456789, 2**32 - 1])
def test_mt19937_matches_numpy(seed):
    raw = (pf.uniform_pos(seed, 2000) * 4294967296.0).astype(np.uint64)
    ref = np.random.RandomState(seed)._bit_generator.random_raw(2000)
    assert np.array_equal(raw, ref)


def test_seed_zero_is_gsl_default():
    assert np.array_equal(pf.uniform_pos(0, 700), pf.uniform_pos(4357, 700))


def test_weibull_inverse_cdf_and_mean():
    n = 200000
    a = np.full(n, 320000.0)
    b = 8.5
    w = pf.weibull(77, a, b)
    u = pf.uniform_pos(77, n)
    # the draws are a (-log u)^(1/b) of the stream's uniforms, in order
    np.testing.assert_allclose(np.exp(-(w / a) ** b), u, rtol=1e-12)
    assert w.mean() == pytest.approx(320000.0 * math.gamma(1 + 1 / b), rel=2e-3)
    # per-voxel scales (k_c_field) scale each draw
    a2 = np.linspace(1.0, 2.0, 1000)
    np.testing.assert_array_equal(pf.weibull(5, a2, b), a2 * pf.weibull(5, np.ones(1000), b))


def test_gb_indicator_checks_minus_y_twice():
    g = np.ones((4, 4, 4), np.int32)
    g[2, 2, 2] = 2
    gb = o.gb_indicator(g)
    # the odd voxel and its +-x, +-y and -z neighbours see another grain
    # through their +x/-x/+y/-y/+z checks; the +z neighbour would need a -z
    # check, which the reference replaces by a second -y check
    assert gb[2, 2, 2] == 1
    for c in ((1, 2, 2), (3, 2, 2), (2, 1, 2), (2, 3, 2), (2, 2, 1)):
        assert gb[c] == 1, c
    assert gb[2, 2, 3] == 0
    assert gb.sum() == 6


def _dens(shape, nsys, value):
    z = np.zeros((nsys,) + shape)
    return z + value, z.copy(), z.copy(), z.copy(), z.copy()


def test_nucleation_threshold_ids_and_spheres():
    shape = (9, 6, 6)
    g = np.ones(shape, np.int32)
    g[5:] = 2
    rs, rm, g1, g2, g3 = _dens(shape, 12, 1.0)
    rm += 0.5
    g1 += 3.0
    g2 += 4.0  # |rho_g| = 5 per system: rho_tot = 12 * 6.5 = 78, diff 6.5
    kappa = np.full(shape, 6.5)          # equal: no nucleus (strict test)
    kappa[4, 1, 1] = 6.0                 # at the boundary plane x = 4|5
    kappa[5, 4, 4] = 6.0
    kappa[5, 0, 0] = 6.0
    kappa[1, 3, 3] = 1.0                 # interior of grain 1: not at a boundary
    r = o.drx_nucleate(g, rs, rm, g1, g2, g3, kappa, 2, 4, 2.0, 3.0, np.zeros(shape), np.ones(shape))
    assert r["status"] == 0 and r["n_new"] == 3
    assert (r["rho_tot"] == 78.0).all()
    # sweep order (4,1,1) < (5,0,0) < (5,4,4): IDs 3, 4, 5
    vox = [np.unravel_index(v, shape) for v in r["vox"]]
    assert vox == [(4, 1, 1), (5, 0, 0), (5, 4, 4)]
    assert r["kc"][4, 1, 1] == 6.0 and r["kc"][1, 3, 3] == 1.0
    # spheres of |d|^2 <= 4: the third nucleus paints its neighbourhood,
    # the second (on the y = 0, z = 0 edge) paints without wrapping to
    # y = 5 or z = 5
    assert r["grain"][5, 4, 4] == 5 and r["grain"][7, 4, 4] == 5 and r["grain"][5, 4, 2] == 5
    assert r["grain"][5, 0, 0] == 4 and r["grain"][5, 0, 2] == 4
    assert r["grain"][5, 0, 5] == 2 and r["grain"][5, 5, 0] == 2
    # the first nucleus lies within the second's sphere (|d|^2 = 3): painted 4
    assert r["grain"][4, 1, 1] == 4 and r["grain"][8, 0, 0] == 2
    # (5, 2, 2) is within reach of both the first (|d|^2 = 3) and the third
    # (|d|^2 = 8, no) nucleus -> 3; (5, 2, 3) reaches the first (|d|^2 = 6, no)
    # and the third (|d|^2 = 5, no) -> unchanged
    assert r["grain"][5, 2, 2] == 3 and r["grain"][5, 2, 3] == 2
    painted = r["grain"] != g
    assert (r["nuclei"][painted] == 1).all() and (r["gid_rex"][painted] == 1).all()
    assert r["nuclei"][~painted].sum() == 0


def test_later_nucleus_paints_over_earlier():
    shape = (6, 6, 6)
    g = np.ones(shape, np.int32)
    g[3:] = 2
    rs, rm, g1, g2, g3 = _dens(shape, 1, 24.0)   # diff_rho = 2
    kappa = np.full(shape, 5.0)
    kappa[2, 2, 2] = 1.0
    kappa[3, 2, 2] = 1.0
    r = o.drx_nucleate(g, rs, rm, g1, g2, g3, kappa, 2, 1, 1.0, 1.0, np.zeros(shape), np.ones(shape))
    assert r["n_new"] == 2
    # both nuclei are within each other's sphere: the later one (3,2,2) wins both
    assert r["grain"][2, 2, 2] == 4 and r["grain"][3, 2, 2] == 4
    assert r["grain"][1, 2, 2] == 3 and r["grain"][4, 2, 2] == 4
