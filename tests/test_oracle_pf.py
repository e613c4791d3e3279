"""CPU checks of the phase-field oracle (oracle/pf_oracle.c, the restatement
of PF_XRX, src/pf_xrx.c:1469-1698, and of PFdata, src/PFgrid.h:54-94).

What pins it:
  * PFdata's lookup on hand-worked lists (present entry, re-label of a last
    entry below 5e-3, append, the empty list, a full list);
  * the 18-point laplacian against the exact Laplacian of quadratics and its
    weights (faces 1/2, edges 1/8, centre -9/2);
  * the coefficients of src/pf_xrx.c:1497-1504 in closed form;
  * invariants of the dynamics: one grain stays one grain; a flat
    bicrystal with equal stored energy keeps its boundary plane and its
    profile depends on x only; with no stored-energy term a
    small grain shrinks, with a strong one a low-energy nucleus grows;
  * the data-parallel model of the GPU kernels (tests/pf_model.py), built on
    four properties of the sweep rather than the sweep itself, agreeing with
    the sequential restatement bit for bit -- on PF_XRX starts, on nuclei
    growth with stored-energy resets, and on random lists that exercise the
    re-label rules.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))

import oracle as o  # noqa: E402
import pf_cases  # noqa: E402
import pf_model  # noqa: E402

EPS = 5e-3


def test_lookup_present_relabel_append():
    assert o.pf_lookup([3, 5], [0.9, 0.1], 5) == ([3, 5], [0.9, 0.1], 1)
    # absent grain, last entry below EPSILON: re-labelled, value kept
    assert o.pf_lookup([3, 5], [0.9, 0.001], 7) == ([3, 7], [0.9, 0.001], 1)
    # absent grain, last entry above: a zero entry is appended
    assert o.pf_lookup([3, 5], [0.9, 0.1], 7) == ([3, 5, 7], [0.9, 0.1, 0.0], 2)
    # a small entry that is not last is not re-labelled
    assert o.pf_lookup([5, 3], [0.001, 0.9], 7) == ([5, 3, 7], [0.001, 0.9, 0.0], 2)
    assert o.pf_lookup([], [], 4) == ([4], [0.0], 0)
    # exactly EPSILON is not below it
    assert o.pf_lookup([3], [EPS], 7) == ([3, 7], [EPS, 0.0], 1)
    # a full list that must append reports -1
    assert o.pf_lookup([1, 2], [0.5, 0.5], 9, cap=2)[2] == -1


def test_lookup_sequence_appends_once():
    idx, val = [4], [1.0]
    for g in (6, 8, 6, 9):
        idx, val, _ = o.pf_lookup(idx, val, g)
    # one zero entry, re-labelled by every later absent grain
    assert idx == [4, 9] and val == [1.0, 0.0]


def test_laplacian_exact_on_quadratics():
    offs = pf_model.LAPN
    for f, lap in ((lambda x, y, z: x * x, 2.0), (lambda x, y, z: y * y, 2.0),
                   (lambda x, y, z: z * z, 2.0), (lambda x, y, z: x * y, 0.0),
                   (lambda x, y, z: y * z - 3 * x + 7.0, 0.0),
                   (lambda x, y, z: 0.5 * x * x + y * y - z * z, 1.0)):
        p = np.array([2.0, -1.0, 3.0])
        a = [f(*(p + np.array(o_))) for o_ in offs] + [f(*p)]
        assert o.pf_laplacian(a) == pytest.approx(lap, abs=1e-12)


def test_laplacian_weights():
    for k in range(19):
        a = np.zeros(19)
        a[k] = 1.0
        want = 0.5 if k < 6 else (0.125 if k < 18 else -4.5)
        assert o.pf_laplacian(a) == want


def test_coefficients_closed_form():
    P = pf_cases.params(8)
    c = o.pf_coeffs(P)
    dx, dt = P["pf_length_scale"], P["pf_time_step"]
    L = 3 * dx
    np.testing.assert_allclose(c, [dt * 4 / 3 * P["M_gb"] / L, 0.75 * P["E_gb"] * L / dx**2,
                                   6 * P["E_gb"] / L, P["alpha"] * P["mu"] * P["bb"] ** 2], rtol=1e-15)


def test_single_grain_is_stationary():
    gid = np.full((6, 5, 4), 3, np.int32)
    rho = np.full(gid.shape, 150.0)
    r = o.pf_xrx(gid, rho, 5, pf_cases.params(3), np.zeros(10))
    assert r["status"] == 0
    assert np.array_equal(r["gid"], gid)
    assert (r["n"] == 1).all() and (r["idx"][..., 0] == 3).all() and (r["val"][..., 0] == 1.0).all()
    # no voxel holds two grains, so nothing writes rho_avg
    assert not r["rho_avg"].any()


def test_flat_bicrystal_keeps_its_plane():
    L = 12
    gid = np.ones((L, 6, 5), np.int32)
    gid[L // 2:] = 2
    rho = np.full(gid.shape, 1.0)  # equal, exactly representable averages
    r = o.pf_xrx(gid, rho, 12, pf_cases.params(2), np.zeros(4))
    assert r["status"] == 0
    assert np.array_equal(r["gid"], gid)
    # eta_1 along x, the same in every (y, z) column
    eta1 = np.where(r["idx"] == 1, r["val"], 0.0).sum(-1)
    assert (eta1 == eta1[:, :1, :1]).all()
    prof = eta1[:, 0, 0]
    # grain 1 fills x < 6 away from the boundaries, and its order parameter
    # falls towards the boundary at 5|6 (the x-major lookup order makes the
    # two sides of a boundary differ, so the profile is not mirror-symmetric)
    assert prof[2] > 0.99 and prof[3] > 0.99
    assert all(prof[k] >= prof[k + 1] for k in range(2, L // 2 - 1))
    assert 0.5 < prof[L // 2 - 1] < 1.0
    assert ((r["val"] >= 0.0) & (r["val"] <= 1.0 + 1e-12)).all()


def _grain_volume(r, g):
    return int((r["gid"] == g).sum())


def test_curvature_shrinks_and_stored_energy_grows_a_nucleus():
    gid = np.ones((14, 14, 14), np.int32)
    idx = np.indices(gid.shape).reshape(3, -1).T
    m = (((idx - 7) ** 2).sum(1) <= 5).reshape(gid.shape)
    gid[m] = 2
    rho = np.where(m, 96.0, 250.0)
    v0 = int(m.sum())
    shrink = o.pf_xrx(gid, rho, 30, pf_cases.params(1, alpha=0.0), np.zeros(4))
    grow = o.pf_xrx(gid, rho, 30, pf_cases.params(1, bb=1.0), np.zeros(4))
    assert shrink["status"] == 0 and grow["status"] == 0
    assert _grain_volume(shrink, 2) < v0 < _grain_volume(grow, 2)
    # the growing nucleus resets the stored energy it sweeps to 100
    assert (grow["defeng"] == 100.0).sum() > 0 and (shrink["defeng"] == 100.0).sum() == 0


def _same(model, r):
    gid_new, A, de, ra = model
    shape = r["gid"].shape
    assert r["status"] == 0
    assert np.array_equal(gid_new, r["gid"])
    assert np.array_equal(np.array(ra), r["rho_avg"])
    for x in range(shape[0]):
        for y in range(shape[1]):
            for z in range(shape[2]):
                n = r["n"][x, y, z]
                ref = [[int(r["idx"][x, y, z, k]), float(r["val"][x, y, z, k])] for k in range(n)]
                assert A[(x + 3, y + 3, z + 3)] == ref, (x, y, z)
                assert de[(x + 3, y + 3, z + 3)] == r["defeng"][x, y, z]


@pytest.mark.parametrize("seed", [1, 2])
def test_parallel_model_matches_sweep_pf_xrx(seed):
    gid, rho, P = pf_cases.case((10, 9, 8), 6, 3, seed)
    ra = np.zeros(40)
    r = o.pf_xrx(gid, rho, 6, P, ra)
    _same(pf_model.run(gid, rho, 6, P, ra), r)


def test_parallel_model_matches_sweep_with_resets():
    gid, rho, P = pf_cases.case((12, 12, 12), 6, 3, 11)
    ra = np.zeros(40)
    r = o.pf_xrx(gid, rho, 10, P, ra)
    assert (r["defeng"] == 100.0).sum() > 10
    _same(pf_model.run(gid, rho, 10, P, ra), r)


@pytest.mark.parametrize("seed", [3, 4])
def test_parallel_model_matches_sweep_random_lists(seed):
    gid, rho, P = pf_cases.case((9, 8, 7), 7, 2, seed)
    n, idx, val = pf_cases.random_lists(gid, seed + 100, cap=o.PF_CAP)
    ra = np.random.default_rng(seed).uniform(0, 300, 40)
    ra[::5] = 0.0
    r = o.pf_run_lists(n, idx, val, gid, rho, 3, P, ra)
    _same(pf_model.run_lists(n, idx, val, gid, rho, 3, P, ra), r)


def test_unstable_step_trips_the_sum_check():
    # ten times the mobility destabilises the explicit step: the order
    # parameters of some voxel fall below a sum of 0.5, where the reference
    # asserts (src/pf_xrx.c:1383-1386)
    gid = np.ones((14, 14, 14), np.int32)
    idx = np.indices(gid.shape).reshape(3, -1).T
    m = (((idx - 7) ** 2).sum(1) <= 5).reshape(gid.shape)
    gid[m] = 2
    rho = np.where(m, 96.0, 250.0)
    r = o.pf_xrx(gid, rho, 30, pf_cases.params(1, alpha=0.0, M_gb=7.4e-10), np.zeros(4))
    assert r["status"] == o.PF_ERR_SUM


def test_grain_outside_rho_avg_is_an_error():
    gid = np.ones((5, 5, 5), np.int32)
    gid[2:] = 9
    r = o.pf_xrx(gid, np.full(gid.shape, 10.0), 2, pf_cases.params(1), np.zeros(4))
    assert r["status"] == o.PF_ERR_GRAIN
