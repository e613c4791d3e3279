"""GPU parity of the phase-field recrystallisation step (PF_XRX,
src/pf_xrx.c:1469-1698) with the oracle's sequential restatement
(oracle/pf_oracle.c): grain field (Update_gID), every order-parameter list
entry (grain and value), the stored-energy buffer and rho_avg, all bit for
bit, through the evpgpu_pf_* C ABI."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
sys.path.insert(0, os.path.join(HERE, "..", "m-themes_amd"))

import oracle as o  # noqa: E402
import pf_cases  # noqa: E402
from pyevp import lib  # noqa: E402
from pyevp import pf as gpf  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def same(g, r, gid_new, ra):
    """GPU state g (lists()) and outputs against the oracle result r"""
    assert r["status"] == 0
    assert np.array_equal(gid_new, r["gid"])
    assert np.array_equal(ra, r["rho_avg"])
    assert np.array_equal(g["n"], r["n"])
    k = gpf.CAP
    n = g["n"][..., None]
    mask = np.arange(k)[None, None, None, :] < n
    assert np.array_equal(np.where(mask, g["idx"], 0), np.where(mask, r["idx"][..., :k], 0))
    assert np.array_equal(np.where(mask, g["val"], 0.0), np.where(mask, r["val"][..., :k], 0.0))
    assert np.array_equal(g["defeng"], r["defeng"])


def run_both(gid, rho, P, steps, n_rho=64, ra=None):
    ra = np.zeros(n_rho) if ra is None else ra
    say(f"pf: shape {gid.shape}, {steps} substeps, grains {int(gid.max())}")
    r = o.pf_xrx(gid, rho, steps, P, ra)
    pf = gpf.PhaseField(gid.shape, ra.size)
    try:
        gn, ra_g = pf.pf_xrx(gid, rho, steps, P, ra)
        st = pf.stats()
        say(f"pf: stats {st}")
        return r, gn, ra_g, pf.lists(), st
    finally:
        pf.close()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pf_xrx_small_vs_oracle(seed):
    gid, rho, P = pf_cases.case((10, 9, 8), 6, 3, seed)
    r, gn, ra, g, _ = run_both(gid, rho, P, 6)
    same(g, r, gn, ra)


def test_pf_xrx_growth_with_resets_vs_oracle():
    gid, rho, P = pf_cases.case((16, 16, 16), 8, 5, 11)
    r, gn, ra, g, st = run_both(gid, rho, P, 10)
    assert (r["defeng"] == 100.0).sum() > 10 and (r["gid"] != gid).sum() > 10
    assert st["trigger_passes"] > st["substeps"]  # the resets took extra passes
    same(g, r, gn, ra)


@pytest.mark.parametrize("shape", [(3, 3, 3), (7, 5, 3), (3, 8, 5), (12, 4, 9)])
def test_pf_xrx_shapes_vs_oracle(shape):
    gid, rho, P = pf_cases.case(shape, 4, 1, 7, radius2=1)
    r, gn, ra, g, _ = run_both(gid, rho, P, 4)
    same(g, r, gn, ra)


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_pf_random_lists_vs_oracle(seed):
    """starts whose lists carry entries below 5e-3 in every position (the
    re-label rules of the lookup) and a non-zero rho_avg"""
    gid, rho, P = pf_cases.case((9, 8, 7), 7, 2, seed)
    n, idx, val = pf_cases.random_lists(gid, seed + 100, cap=o.PF_CAP)
    ra0 = np.random.default_rng(seed).uniform(0, 300, 40)
    ra0[::5] = 0.0
    r = o.pf_run_lists(n, idx, val, gid, rho, 3, P, ra0)
    pf = gpf.PhaseField(gid.shape, ra0.size)
    try:
        pf.set_lists(n, idx[..., :gpf.CAP], val[..., :gpf.CAP], gid, rho)
        gn, ra = pf.steps(P, 3, ra0)
        same(pf.lists(), r, gn, ra)
    finally:
        pf.close()


def test_pf_errors():
    # an unstable step: the reference's sum assert
    gid = np.ones((14, 14, 14), np.int32)
    idx = np.indices(gid.shape).reshape(3, -1).T
    m = (((idx - 7) ** 2).sum(1) <= 5).reshape(gid.shape)
    gid[m] = 2
    rho = np.where(m, 96.0, 250.0)
    pf = gpf.PhaseField(gid.shape, 4)
    try:
        with pytest.raises(lib.EvpError) as e:
            pf.pf_xrx(gid, rho, 30, pf_cases.params(1, alpha=0.0, M_gb=7.4e-10), np.zeros(4))
        assert e.value.code == gpf.ERR_PF_SUM
        # a grain beyond rho_avg
        g2 = np.ones((14, 14, 14), np.int32)
        g2[5:] = 9
        with pytest.raises(lib.EvpError) as e:
            pf.pf_xrx(g2, rho, 2, pf_cases.params(1), np.zeros(4))
        assert e.value.code == gpf.ERR_PF_GRAIN
        # the context is usable again
        gn, _ = pf.pf_xrx(gid, rho, 2, pf_cases.params(1), np.zeros(4))
        assert (gn >= 1).all()
    finally:
        pf.close()
    # 27 grains around one voxel: beyond the list capacity
    g3 = np.arange(1, 28, dtype=np.int32).reshape(3, 3, 3)
    pf = gpf.PhaseField(g3.shape, 64)
    try:
        with pytest.raises(lib.EvpError) as e:
            pf.pf_xrx(g3, np.full(g3.shape, 50.0), 1, pf_cases.params(27), np.zeros(64))
        assert e.value.code == gpf.ERR_PF_CAP
    finally:
        pf.close()


def test_pf_xrx_128_vs_oracle():
    """the DRX configuration's grid: 128^3, 256 grains and 40 nuclei, six
    substeps against the oracle"""
    gid, rho, P = pf_cases.case((128, 128, 128), 256, 40, 5, radius2=5)
    r, gn, ra, g, st = run_both(gid, rho, P, 6, n_rho=60000)
    say(f"pf 128: changed voxels {(r['gid'] != gid).sum()}, resets {(r['defeng'] == 100).sum()}, "
        f"{st['ms'] / 6:.2f} ms per substep")
    same(g, r, gn, ra)


def nucleation_inputs(shape, grains, seed, nsys=12):
    gid, _, _ = pf_cases.case(shape, grains, 0, seed)
    rng = np.random.default_rng(seed)
    dens = [rng.uniform(0.5, 3.0, (nsys,) + shape) for _ in range(2)]
    dens += [rng.normal(0.0, 2.0, (nsys,) + shape) for _ in range(3)]
    # static strengths: the reference's Weibull draws (k_c 320000 scaled to
    # these densities, alpha 8.5) from a fixed seed; a few per cent of the
    # boundary voxels nucleate
    kappa = gpf.weibull(seed, np.full(shape, 9.0), 8.5)
    return gid, dens, kappa


@pytest.mark.parametrize("shape,grains,seed", [((24, 20, 16), 12, 1), ((128, 128, 128), 256, 2)])
def test_drx_nucleate_vs_oracle(shape, grains, seed):
    gid, dens, kappa = nucleation_inputs(shape, grains, seed)
    rex0 = np.zeros(shape, np.int32)
    kc0 = np.full(shape, 5.0)
    r = o.drx_nucleate(gid, *dens, kappa, grains, 4, 1.0, 5.0, rex0, kc0, max_new=1 << 20)
    g = gpf.nucleate(gid, *dens, kappa, grains, 4, 1.0, 5.0, rex0, kc0, max_new=1 << 20)
    say(f"nucleation {shape}: {r['n_new']} nuclei, {int((r['grain'] != gid).sum())} voxels repainted")
    assert r["status"] == 0 and r["n_new"] > 0
    assert g["n_new"] == r["n_new"] and np.array_equal(g["vox"], r["vox"])
    for k in ("grain", "gid_rex", "nuclei", "kc", "rho_tot"):
        assert np.array_equal(g[k], r[k]), k


def test_drx_event_pieces_on_a_dislocation_density_state(tmp_path):
    """The event's pieces chained on the state of a real dislocation-density
    run, GPU against oracle at each step: the strengths (seeded Weibull
    draws), the nucleation check on the run's densities and grains, the
    stored energy PF_XRX starts from (8 per slip system at the nuclei, the
    density total elsewhere, src/evolution.c:973-1000) and ten phase-field
    substeps; the grain field that comes out is the same bit for bit."""
    from pyevp import synth
    cfg = synth.Config(n=24, grains=16, seed=0x5EED0D, increments=3, dd=synth.DDLaw())
    job = lib.GpuJob(synth.write_job(str(tmp_path), cfg), lib.FP64, law=lib.LAW_DD)
    job.run(3)
    dd = job.dd_state()
    grain, _ = job.ids()
    n_grains = int(len(job.grain_list()))
    job.close()
    shape = (cfg.n,) * 3
    dens = [np.ascontiguousarray(dd[k].T).reshape((12,) + shape) for k in ("rho_s", "rho_m", "rho_g1", "rho_g2",
                                                                        "rho_g3")]
    g0 = grain.reshape(shape)
    # strengths scaled so that a handful of boundary voxels nucleate (a
    # Weibull scale 5 % above the largest rho_tot / 12)
    tot = sum(d for d in dens[:2]).sum(0) + np.sqrt(dens[2] ** 2 + dens[3] ** 2 + dens[4] ** 2).sum(0)
    kappa = gpf.weibull(0x5EED, np.full(shape, 1.05 * float((tot / 12).max())), 8.5)
    rex0 = np.zeros(shape, np.int32)
    kc0 = np.full(shape, 1.0)
    r = o.drx_nucleate(g0, *dens, kappa, n_grains, 4, 1.0, 1.0, rex0, kc0)
    g = gpf.nucleate(g0, *dens, kappa, n_grains, 4, 1.0, 1.0, rex0, kc0)
    say(f"drx event: {r['n_new']} nuclei, {int((r['grain'] != g0).sum())} voxels repainted")
    assert r["status"] == 0 and r["n_new"] > 0
    for k in ("grain", "nuclei", "gid_rex", "kc", "rho_tot"):
        assert np.array_equal(g[k], r[k]), k
    rho_pf = np.where(g["nuclei"] == 1, 8.0 * 12, g["rho_tot"])
    P = pf_cases.params(n_grains)
    ra = np.zeros(60000)
    ro = o.pf_xrx(g["grain"], rho_pf, 10, P, ra)
    pf = gpf.PhaseField(shape, ra.size)
    try:
        gn, ra_g = pf.pf_xrx(g["grain"], rho_pf, 10, P, ra)
        same(pf.lists(), ro, gn, ra_g)
    finally:
        pf.close()
    say(f"drx event: grains changed by the phase field {int((gn != g['grain']).sum())}")
