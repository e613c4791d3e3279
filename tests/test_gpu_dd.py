"""The dislocation-density law on the GPU (StrainRate_Orowan in the Newton
kernel, the per-increment constants, the shear-rate gradient and
DislocationEvolution) against the oracle, through the C ABI."""
import dataclasses

import numpy as np
import pytest

import oracle as O
from pyevp import lib, synth

pytestmark = pytest.mark.gpu

DD_FIELDS = (("rho_s", O.RHO_S), ("rho_m", O.RHO_M), ("rho_g1", O.RHO_G1), ("rho_g2", O.RHO_G2),
             ("rho_g3", O.RHO_G3), ("rho_F", O.RHO_F), ("rho_P", O.RHO_P))


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _job(tmp_path, name, **kw):
    law = kw.pop("law", synth.DDLaw())
    cfg = synth.Config(dd=law, **kw)
    return synth.write_job(str(tmp_path / name), cfg)


def _compare(J, g, tol=1e-10, dens_tol=1e-10, avg_tol=1e-10):
    s = g.state()
    assert rel(s["Sig"], J.field(O.SIG)) < tol, rel(s["Sig"], J.field(O.SIG))
    assert rel(s["Edot"], J.field(O.EDOT)) < tol
    assert rel(s["Eps"], J.field(O.EPS)) < tol
    assert rel(s["DisGrad"], J.field(O.DISGRAD)) < tol
    assert rel(s["TranMat"], J.field(O.TRANMAT)) < tol
    assert rel(s["gamdot"], J.field(O.GAMDOT)) < tol
    d = g.dd_state()
    for k, f in DD_FIELDS:
        ref = J.field(f)
        if np.abs(ref).max() == 0.0:
            assert np.abs(d[k]).max() == 0.0, k
        else:
            assert rel(d[k], ref) < dens_tol, (k, rel(d[k], ref))
    mo, mg = J.macro(), g.macro()
    assert rel(mg["sig_avg"], mo["SigAvg"]) < tol
    # AvgDislocation (the dd.out columns), dipole and forest sums included
    np.testing.assert_allclose(g.dd_averages(), mo["dd_avg"], rtol=avg_tol, atol=0)


def test_dd_loop_parity_8(tmp_path):
    """Three increments, so the third forms the shear-rate gradient (Hard_Flag,
    istep > 1) and the GND densities move."""
    inp = _job(tmp_path, "dd8", n=8, grains=8, seed=11, increments=3)
    J = O.Job(inp, O.LAW_DD)
    g = lib.GpuJob(inp, lib.FP64, law=lib.LAW_DD)
    it_o = J.run(3)
    it_g, _ = g.run(3)
    assert list(it_g) == list(it_o), (it_g, it_o)
    assert np.abs(J.field(O.RHO_G1)).max() > 0
    _compare(J, g)
    g.close()


def test_dd_loop_parity_16_pores(tmp_path):
    inp = _job(tmp_path, "dd16", n=16, grains=8, seed=12, increments=3, pore_fraction=0.03,
               pore_radius=2)
    J = O.Job(inp, O.LAW_DD)
    g = lib.GpuJob(inp, lib.FP64, law=lib.LAW_DD)
    it_o = J.run(3)
    it_g, _ = g.run(3)
    assert list(it_g) == list(it_o), (it_g, it_o)
    _compare(J, g)
    g.close()


def test_dd_newton_shortcut_is_exact(tmp_path):
    inp = _job(tmp_path, "dd8s", n=8, grains=8, seed=13, increments=3)
    a = lib.GpuJob(inp, lib.FP64, law=lib.LAW_DD)
    b = lib.GpuJob(inp, lib.FP64, law=lib.LAW_DD)
    b.set_newton_shortcut(False)
    ia, _ = a.run(3)
    ib, _ = b.run(3)
    assert list(ia) == list(ib)
    sa, sb = a.state(), b.state()
    for k in ("Sig", "Edot", "gamdot", "DisGrad"):
        assert np.array_equal(sa[k], sb[k]), k
    da, db = a.dd_state(), b.dd_state()
    for k in da:
        assert np.array_equal(da[k], db[k]), k
    a.close()
    b.close()


def test_dd_mixed_precision(tmp_path):
    """The mixed loop on the DD law, at the oracle's iteration counts: stress
    within 1e-6."""
    inp = _job(tmp_path, "dd16m", n=16, grains=8, seed=14, increments=2)
    J = O.Job(inp, O.LAW_DD)
    it = J.run(2)
    g = lib.GpuJob(inp, lib.MIXED, law=lib.LAW_DD)
    for k in it:
        g.run(1, forced=int(k))
    s = g.state()
    assert rel(s["Sig"], J.field(O.SIG)) < 1e-6, rel(s["Sig"], J.field(O.SIG))
    d = g.dd_state()
    assert rel(d["rho_s"], J.field(O.RHO_S)) < 1e-6
    g.close()


def test_dd_without_tau_drx_freezes_the_first_increment(tmp_path):
    """tau_DRX absent (0): the DRX term of the passing stress is 0 * NaN at
    TimeTot = 0 (src/constitutive.c:585), so nothing slips in the first
    increment, and -0/0 turns into exp(-inf) = 0 afterwards."""
    law = synth.DDLaw(tau_drx=None)
    inp = _job(tmp_path, "dd8n", n=8, grains=8, seed=15, increments=2, law=law)
    J = O.Job(inp, O.LAW_DD)
    g = lib.GpuJob(inp, lib.FP64, law=lib.LAW_DD)
    it_o = J.run(1)
    it_g, _ = g.run(1)
    assert list(it_g) == list(it_o)
    assert (g.state()["gamdot"] == 0).all() and (J.field(O.GAMDOT) == 0).all()
    it_o = J.run(1)
    it_g, _ = g.run(1)
    assert list(it_g) == list(it_o)
    assert np.abs(J.field(O.GAMDOT)).max() > 0
    _compare(J, g)
    g.close()


def test_dd_reference_test_run_constants_freeze_slip(tmp_path):
    """Q_slip = 2e10 (the reference's test_run input): exp(-q_slip) = 0, no slip."""
    law = synth.DDLaw(q_slip=2.0e10, T_K=30.0)
    inp = _job(tmp_path, "dd8f", n=8, grains=8, seed=16, increments=2, law=law)
    g = lib.GpuJob(inp, lib.FP64, law=lib.LAW_DD)
    g.run(2)
    s = g.state()
    assert (s["gamdot"] == 0).all() and (s["Edot"] == 0).all()
    g.close()


def test_dd_reference_test_run(tmp_path):
    """The reference's own test_run inputs (32^3, 24 % pores) under the law its
    build compiles in -- copper elastic constants, Q_slip = 2e10 (no slip),
    the SSD annihilation term alone moving the densities -- against the
    oracle's fixture (tests/golden/make_golden.py), then through forge with
    EVP_LAW=dd."""
    import os
    import shutil
    import subprocess
    from conftest import GOLDEN
    g = np.load(os.path.join(GOLDEN, "oracle_test_run_dd.npz"))
    inp = os.path.join(GOLDEN, "test_run", "input.in")
    job = lib.GpuJob(inp, lib.FP64, law=lib.LAW_DD)
    iters, sig_avg, vm = [], [], []
    for _ in range(2):
        it, _ = job.run(1)
        iters.append(int(it[0]))
        m = job.macro()
        sig_avg.append(m["sig_avg"].ravel())
        vm.append([m[k] for k in ("time_tot", "e_vm", "evmp", "d_vm", "dvmp", "s_vm", "s_vm1")])
    assert iters == list(g["iters"]), (iters, list(g["iters"]))
    assert rel(np.array(sig_avg), g["sig_avg"]) < 1e-10
    assert rel(np.array(vm), g["vm"]) < 1e-10
    s = job.state()
    sel = g["sel"]
    assert (s["gamdot"] == 0).all()
    for k, f in (("Sig", "sig"), ("Edot", "edot"), ("DisGrad", "disgrad"), ("TranMat", "tranmat")):
        if np.abs(g[f]).max() == 0.0:
            assert np.abs(s[k][sel]).max() == 0.0, k
        else:
            assert rel(s[k][sel], g[f]) < 1e-10, k
    d = job.dd_state()
    for k in ("rho_s", "rho_m", "rho_F", "rho_P"):
        assert rel(d[k][sel], g[k]) < 1e-10, k
    _, ph = job.ids()
    assert (d["rho_s"][ph == 1] < 0.05).all()
    job.close()
    for f in ("input.in", "Cu_p.sx", "FFT_microstructure_pore_197.txt"):
        shutil.copy(os.path.join(GOLDEN, "test_run", f), tmp_path / f)
    env = dict(os.environ, EVP_LAW="dd")
    r = subprocess.run([os.path.join(lib.LIBDIR, "forge"), "input.in"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    rows = open(tmp_path / "vm.out").read().splitlines()
    vals = np.array([[float(x) for x in ln.split("\t")] for ln in rows[1:]])
    # vm.out carries %e (7 significant digits)
    assert np.allclose(vals, g["vm"], rtol=1e-6, atol=0)
    # dd.out (src/init.c:498, src/evolution.c:1565-1570): the header, then
    # TimeTot, e_vm and AvgDislocation, one row per increment; drx.out keeps
    # its header only (no DRX event)
    lines = open(tmp_path / "dd.out").read().splitlines()
    assert lines[0] == "TIME\tEVM\tDD_F\tDD_P\tDD_s\tDD_m\tDD_g"
    dd = np.array([[float(x) for x in ln.split("\t")] for ln in lines[1:]])
    assert open(tmp_path / "drx.out").read() == "TIME\tEVM\tNEW_IDs\n"
    assert dd.shape == (2, 9)
    np.testing.assert_allclose(dd[:, :2], g["vm"][:, :2], rtol=1e-6, atol=0)
    np.testing.assert_allclose(dd[-1, 2:], g["dd_avg"], rtol=1e-6, atol=0)


def test_dd_loop_parity_simple_shear(tmp_path):
    """Fully strain-controlled simple shear under the DD law: lattice spin
    from the 12 shear rate, so update_orient turns the orientations between
    the densities' gradient steps."""
    inp = _job(tmp_path, "dd8sh", n=8, grains=8, seed=17, increments=3,
               velgrad_bc_flag=(1, 1, 1, 1, 1, 1), velgrad_bc=(0.0, 0.0, 0.0, 0.0, 0.0, 1.0),
               stress_bc_flag=(0, 0, 0, 0, 0, 0))
    J = O.Job(inp, O.LAW_DD)
    g = lib.GpuJob(inp, lib.FP64, law=lib.LAW_DD)
    it_o = J.run(3)
    it_g, _ = g.run(3)
    assert list(it_g) == list(it_o), (it_g, it_o)
    _compare(J, g)
    g.close()


def test_dd_state_calls_refused_where_they_do_not_apply(tmp_path):
    """The reverse pass is of the power law: a DD context refuses it; the
    density calls refuse a power-law context (EVPGPU_ERR_STATE, no crash)."""
    inp = _job(tmp_path, "dd8r", n=8, grains=4, seed=21, increments=2)
    g = lib.GpuJob(inp, lib.FP64, law=lib.LAW_DD)
    with pytest.raises(lib.EvpError) as e:
        g.adjoint_record(2, 1)
    assert e.value.code == -5
    g.close()
    p = lib.GpuJob(inp, lib.FP64)
    for call in (p.dd_averages, p.dd_state):
        with pytest.raises(lib.EvpError) as e:
            call()
        assert e.value.code == -5
    p.close()


def test_dd_loop_parity_non_cubic_any_size(tmp_path):
    """12 x 10 x 9 with pores: the shear-rate gradient's neighbour indexing
    on a grid whose three lengths differ (the reference's grain check on the
    z-minus side reads the -y neighbour, src/constitutive.c:818,838,858) and
    the two-level DFT transforms of lengths 10 and 9; three increments, so
    the GND densities move."""
    inp = _job(tmp_path, "ddany", n=0, dims=(12, 10, 9), grains=6, seed=61, increments=3,
               pore_fraction=0.02, pore_radius=1)
    J = O.Job(inp, O.LAW_DD)
    g = lib.GpuJob(inp, lib.FP64, law=lib.LAW_DD)
    it_o = J.run(3)
    it_g, _ = g.run(3)
    assert list(it_g) == list(it_o), (it_g, it_o)
    assert np.abs(J.field(O.RHO_G1)).max() > 0
    _compare(J, g)
    g.close()


def test_dd_loop_parity_full_size_128(tmp_path):
    """The 128^3 job of the benchmark (512 grains, 2 % pores) under the DD
    law at full size: increment 1 forced to 2 Green iterations in the oracle
    (about a minute of CPU) and on the GPU, then the densities' evolution;
    FP64 to the loop bar on every field, the mixed loop to 1e-6."""
    inp = synth.write_job(str(tmp_path), dataclasses.replace(synth.config_c2(), dd=synth.DDLaw()))
    J = O.Job(inp, O.LAW_DD)
    J.run(1, forced=[2])
    g = lib.GpuJob(inp, lib.FP64, law=lib.LAW_DD)
    g.run(1, forced=2)
    # the oracle's averages are sums of 12 x 2^21 positive terms taken in
    # sequence: their rounding bound is n u = 12 x 2^21 x 2^-53 = 2.8e-9 of
    # the sum (the GPU sums in a tree); the fields themselves are held to the
    # loop bar
    _compare(J, g, avg_tol=3e-9)
    g.close()
    g = lib.GpuJob(inp, lib.MIXED, law=lib.LAW_DD)
    g.run(1, forced=2)
    assert rel(g.state()["Sig"], J.field(O.SIG)) < 1e-6, rel(g.state()["Sig"], J.field(O.SIG))
    d = g.dd_state()
    for k, f in DD_FIELDS[:2]:
        assert rel(d[k], J.field(f)) < 1e-6, (k, rel(d[k], J.field(f)))
    g.close()
