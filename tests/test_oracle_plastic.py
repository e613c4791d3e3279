"""Pins of the oracle's plastic whole loop against exact solutions.

A uniform single crystal (no pores) and a two-crystal laminate along x are
problems whose converged fields the mechanics fixes without any Green
operator: the fields are uniform per crystal, equilibrium is the continuity
of the traction sigma . e_x, compatibility allows strain jumps sym(c x e_x)
only, the macroscopic mixed conditions hold, and each crystal satisfies the
implicit elasto-viscoplastic update S sigma + eps_p + dt edot(sigma) = E.
tests/crystal_np.py solves exactly that system (nested Newton solves in
numpy) and then applies the increment's updates: plastic strain, the
lattice rotation by the exponential of the spin, the Voce hardening in
closed form.

The oracle runs the reference's loop (src/evolution.c:1318-1547, Q1
reconstruction) on the same inputs in the reference's file formats, with
the loop tolerance at 1e-12 and the validation Newton tolerances; every
increment's stress and displacement gradient and the final plastic strain,
orientation, CRSS and accumulated slip must agree.  Slips a plausible
restatement error would cause (a wrong increment counter, a transposed
rotation, a missing skew(L) term, a Voce sign, a mixed-BC index) each fail
here; the first made this test fail at 3e-2 while it was written.
"""
import os

import numpy as np
import pytest

import oracle as O
from pyevp import synth
import crystal_np as X

K = 10


def write(d, eulers, mat):
    dims = (2, 2, 2) if len(eulers) == 1 else (4, 2, 2)
    cfg = synth.Config(n=0, dims=dims, grains=len(eulers), increments=K, err=1e-12, iter_max=40000,
                       material=mat)
    N = int(np.prod(dims))
    x = np.arange(N) // (dims[1] * dims[2])
    lay = x < dims[0] // 2 if len(eulers) == 2 else np.ones(N, bool)
    eul = np.where(lay[:, None], np.array(eulers[0]), np.array(eulers[-1]))
    grain = np.where(lay, 1, 2).astype(np.int32)
    synth.write_ms(os.path.join(d, "ms.txt"), dims, eul, grain, np.ones(N, np.int32))
    synth.write_sx(os.path.join(d, "Cu.sx"), mat)
    synth.write_input(os.path.join(d, "input.in"), cfg)
    return os.path.join(d, "input.in"), lay


@pytest.fixture
def tight_newton():
    O.set_newton_tol(1e-14, 1e-13)
    yield
    O.set_newton_tol()


def run_both(tmp_path, eulers, mat, tol):
    inp, lay = write(str(tmp_path), eulers, mat)
    J = O.Job(inp)
    M = X.Laminate(eulers, mat)
    for k in range(K):
        J.run(1)
        sigs, res = M.increment()
        assert res < 1e-9
        ref = np.where(lay[:, None, None], sigs[0][None], sigs[-1][None])
        s = J.field(O.SIG)
        assert np.max(np.abs(s - ref)) <= tol * np.max(np.abs(ref)), (k, np.max(np.abs(s - ref)))
        grad = np.where(lay[:, None, None], M.L[0].grad[None], M.L[-1].grad[None])
        dg = J.field(O.DISGRAD)
        assert np.max(np.abs(dg - grad)) <= tol * np.max(np.abs(grad)), k
        m = J.macro()["SigAvg"]
        assert np.max(np.abs(m - sum(sigs) / len(sigs))) <= tol * np.max(np.abs(ref))
    tm, cr, ep, ga = (J.field(f) for f in (O.TRANMAT, O.CRSS, O.EPS, O.GAMACUM))
    for l, L in enumerate(M.L):
        v = 0 if l == 0 else len(lay) - 1
        assert np.max(np.abs(tm[v] - L.T)) <= tol
        assert np.max(np.abs(cr[v] - L.crss[:, None])) <= tol * np.max(L.crss)
        assert np.max(np.abs(ep[v] - L.epsp)) <= tol * np.max(np.abs(L.epsp))
        assert abs(ga[v] - L.G) <= tol * L.G
        assert L.G > 0 and np.max(L.crss) > mat.tau0f  # hardening happened
    return M


@pytest.mark.parametrize("euler", [(20.0, 35.0, 50.0), (135.0, 60.0, 10.0), (0.0, 0.0, 0.0)])
def test_uniform_single_crystal_matches_single_point_solution(tmp_path, tight_newton, euler):
    M = run_both(tmp_path, [euler], synth.Material(), 1e-10)
    # the lattice rotated (orientation update on) unless the crystal is a
    # symmetric cube orientation under [001] tension
    if euler != (0.0, 0.0, 0.0):
        assert np.max(np.abs(M.L[0].T - X.rotation(euler))) > 1e-6


@pytest.mark.parametrize("pair", [((10.0, 30.0, 40.0), (70.0, 20.0, 110.0)),
                                  ((250.0, 80.0, 300.0), (45.0, 54.7356, 0.0))])
def test_plastic_laminate_matches_exact_laminate_solution(tmp_path, tight_newton, pair):
    mat = synth.Material(nrs=12.0, tau1=20.0, thet0=300.0)
    run_both(tmp_path, list(pair), mat, 1e-9)
