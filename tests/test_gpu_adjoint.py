"""Hand-written adjoint of the loop (evpgpu_adjoint_record / evpgpu_adjoint_run).

The job is the one of tests/golden/make_adjoint_golden.py: 8^3 voxels,
8 grains, Cu, uniaxial tension, 5 increments (orientation update and
hardening from the third on).  Forward runs use the validation Newton
tolerances (step 1e-14, TOLX 1e-13) and Green tolerance 1e-11, so the loop
output is smooth in the parameters at the 1e-6 finite-difference scale;
with the reference's tolerances TOLX quantises it (last test).

Checks:
  * the gradient against the FP64 reference gradient of the oracle (central
    differences of the independent CPU restatement, the committed fixture);
  * the gradient against central differences (step 1e-6 relative) of the
    GPU forward itself, each perturbed run forced to the base iteration counts;
  * rows of the Jacobian d Sigma_k / d p, checkpoint-interval invariance,
    the torch.autograd.Function wrapper, the time ratio, the input checks.
"""
import os
import time

import numpy as np
import pytest

from conftest import GOLDEN
from pyevp import lib, synth

pytestmark = pytest.mark.gpu

FIX = os.path.join(GOLDEN, "oracle_adjoint_8.npz")
NINC = 5
TOL_GRAD = 1e-5


def config(p, err=1e-11):
    m = synth.Material(tau0f=p[0], tau0b=p[0], tau1=p[1], thet0=p[2], thet1=p[3], nrs=p[4])
    return synth.Config(n=8, grains=8, seed=0x5EED0008, increments=NINC, err=err, iter_max=40000,
                        material=m)


def open_job(d, p, tight=True, err=1e-11):
    job = lib.GpuJob(synth.write_job(str(d), config(p, err)))
    if tight:
        job.set_newton_tolerances(1e-14, 1e-13)
    return job


def seeds_loss(curve, target):
    """dL/dSigma for L = sum_k (Sigma_33,k - target_k)^2."""
    b = np.zeros((NINC, 6))
    b[:, 2] = 2.0 * (curve[:, 2] - target)
    return b


def forced_curve(job, iters):
    out = []
    for k in range(NINC):
        job.run(1, forced=int(iters[k]))
        s = job.macro()["sig_avg"].reshape(3, 3)
        out.append([s[0, 0], s[1, 1], s[2, 2], s[1, 2], s[0, 2], s[0, 1]])
    return np.array(out)


@pytest.fixture(scope="module")
def fx():
    return np.load(FIX)


@pytest.fixture(scope="module")
def rec(tmp_path_factory, fx):
    """One recorded forward at the base parameters (checkpoints every 2)."""
    job = open_job(tmp_path_factory.mktemp("adj8"), fx["params"])
    t0 = time.perf_counter()
    curve, stats = job.adjoint_record(NINC, checkpoint_every=2)
    t_fwd = time.perf_counter() - t0
    return dict(job=job, curve=curve, stats=stats, t_fwd=t_fwd)


def test_forward_matches_oracle(rec, fx):
    iters = [s["iters"] for s in rec["stats"]]
    assert iters == list(fx["iters"]), (iters, list(fx["iters"]))
    assert np.abs(rec["curve"] - fx["curve"]).max() <= 1e-10 * np.abs(fx["curve"]).max()


def test_gradient_vs_oracle_fd(rec, fx):
    job = rec["job"]
    g, st = job.adjoint_run(seeds_loss(rec["curve"], fx["target"]), cg_tol=1e-12, cg_max=4000)
    gref = fx["grad_fd"]
    print("adjoint", g[0], "\noracle FD", gref, "\nstats", st)
    assert np.all(np.abs(g[0] - gref) <= TOL_GRAD * np.abs(gref)), (g[0], gref)
    assert np.all(g[1] == 0.0)


@pytest.mark.parametrize("k,c", [(4, 2), (2, 2), (4, 5), (3, 3)])
def test_jacobian_rows_vs_oracle_fd(rec, fx, k, c):
    """d Sigma_{k,c} / d p for single components (c: 11 22 33 23 13 12)."""
    job = rec["job"]
    b = np.zeros((NINC, 6))
    b[k, c] = 1.0
    g, _ = job.adjoint_run(b, cg_tol=1e-12, cg_max=4000)
    ref = fx["jac_fd"][k, c]
    # the central differences carry ~1e-10 absolute noise; entries far below
    # the row's largest are compared against 1e-3 of that largest one
    floor = 1e-3 * np.abs(ref).max()
    assert np.all(np.abs(g[0] - ref) <= TOL_GRAD * np.maximum(np.abs(ref), floor)), (g[0], ref)


def test_stress_controlled_components_do_not_move(rec, fx):
    """Sigma_11 and Sigma_22 are prescribed (zero lateral stress): their exact
    derivatives vanish; the finite differences only show the loop tolerance."""
    job = rec["job"]
    big = np.abs(fx["jac_fd"][:, 2]).max(axis=0)
    for c in (0, 1):
        b = np.zeros((NINC, 6))
        b[NINC - 1, c] = 1.0
        g, _ = job.adjoint_run(b, cg_tol=1e-12, cg_max=4000)
        assert np.all(np.abs(g[0]) <= 1e-9 * big), (c, g[0])
        assert np.all(np.abs(fx["jac_fd"][:, c]) <= 1e-6 * big)


def test_gradient_vs_gpu_fd(tmp_path, rec, fx):
    """Central differences (step 1e-6 relative) of the GPU forward itself."""
    p0 = fx["params"]
    iters = [s["iters"] for s in rec["stats"]]
    g, _ = rec["job"].adjoint_run(seeds_loss(rec["curve"], fx["target"]), cg_tol=1e-12, cg_max=4000)
    job = open_job(tmp_path, p0)
    fd = np.zeros(5)
    for i in range(5):
        h = 1e-6 * abs(p0[i])
        lv = []
        for sgn in (1, -1):
            p = p0.copy()
            p[i] += sgn * h
            job.set_params(0, p)
            cv = forced_curve(job, iters)
            lv.append(np.sum((cv[:, 2] - fx["target"]) ** 2))
        fd[i] = (lv[0] - lv[1]) / (2 * h)
    print("adjoint", g[0], "\nGPU FD", fd)
    assert np.all(np.abs(g[0] - fd) <= TOL_GRAD * np.abs(fd)), (g[0], fd)


def test_checkpoint_interval_does_not_change_the_gradient(tmp_path, rec, fx):
    b = seeds_loss(rec["curve"], fx["target"])
    g2, _ = rec["job"].adjoint_run(b, cg_tol=1e-12, cg_max=4000)
    for ck in (1, 5):
        job = open_job(tmp_path / f"ck{ck}", fx["params"])
        curve, _ = job.adjoint_record(NINC, checkpoint_every=ck)
        assert np.array_equal(curve, rec["curve"])
        g, _ = job.adjoint_run(b, cg_tol=1e-12, cg_max=4000)
        assert np.abs(g - g2).max() <= 1e-9 * np.abs(g2).max()
        job.close()


def test_state_after_adjoint_is_the_recorded_final_state(rec):
    job = rec["job"]
    s1 = job.state()
    job.adjoint_run(np.ones((NINC, 6)), cg_tol=1e-10, cg_max=4000)
    s2 = job.state()
    for k in ("Sig", "DisGrad", "TranMat", "crss"):
        assert np.array_equal(s1[k], s2[k]), k


def test_autograd_function(tmp_path, fx):
    torch = pytest.importorskip("torch")
    from pyevp.autograd import MacroStressCurve
    job = open_job(tmp_path, fx["params"])
    p = torch.tensor(fx["params"], dtype=torch.float64, requires_grad=True)
    curve = MacroStressCurve.apply(p, job, NINC, 0, 2, 1e-12, 4000)
    loss = ((curve[:, 2] - torch.from_numpy(fx["target"])) ** 2).sum()
    loss.backward()
    assert abs(loss.item() - float(fx["loss"])) <= 1e-9 * float(fx["loss"])
    g = p.grad.numpy()
    assert np.all(np.abs(g - fx["grad_fd"]) <= TOL_GRAD * np.abs(fx["grad_fd"])), (g, fx["grad_fd"])


def test_adjoint_time_within_three_forwards(rec, fx):
    t0 = time.perf_counter()
    _, st = rec["job"].adjoint_run(seeds_loss(rec["curve"], fx["target"]), cg_tol=1e-10, cg_max=4000)
    t_adj = time.perf_counter() - t0
    print(f"forward {rec['t_fwd']:.3f}s adjoint {t_adj:.3f}s (replay {st['replay_seconds']:.3f}s,"
          f" CG iterations {st['cg_iters']}, max/increment {st['cg_max_iters']})")
    assert t_adj <= 3.0 * rec["t_fwd"]


def test_reference_tolerances_quantise_the_forward(tmp_path, fx):
    """With the reference's Newton tolerances (TOLX 1e-7) the loop output is
    piecewise smooth at the 1e-7 relative level: the adjoint (the derivative
    of the converged equations) still agrees with a central difference at
    step 1e-4 relative to about 1e-3."""
    p0 = fx["params"]
    job = open_job(tmp_path, p0, tight=False, err=1e-8)
    curve, stats = job.adjoint_record(NINC, checkpoint_every=5)
    iters = [s["iters"] for s in stats]
    b = np.zeros((NINC, 6))
    b[NINC - 1, 2] = 1.0
    g, _ = job.adjoint_run(b, cg_tol=1e-12, cg_max=4000)
    fj = open_job(tmp_path / "fd", p0, tight=False, err=1e-8)
    for i in (0, 4):
        h = 1e-4 * abs(p0[i])
        vals = []
        for sgn in (1, -1):
            p = p0.copy()
            p[i] += sgn * h
            fj.set_params(0, p)
            vals.append(forced_curve(fj, iters)[NINC - 1, 2])
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(g[0][i] - fd) <= 2e-3 * abs(fd), (i, g[0][i], fd)


def test_adjoint_refuses_gas_phases(tmp_path):
    cfg = synth.Config(n=8, grains=4, seed=3, increments=1, pore_fraction=0.05, pore_radius=1)
    job = lib.GpuJob(synth.write_job(str(tmp_path), cfg))
    with pytest.raises(lib.EvpError):
        job.adjoint_record(1, 1)
