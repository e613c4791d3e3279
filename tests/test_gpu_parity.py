"""Parity of the HIP path (through the C ABI) against the oracle and numpy.

Tolerances (FP64 path): FFTs 1e-12 of the spectrum's max; Green operator 1e-12;
per-voxel Newton 1e-12 on stress with equal iteration counts; whole loop 1e-10
normwise-relative on stress, strain rate and the macroscopic curve, with equal
Green iteration counts per increment.  Mixed path: 1e-6 on stress.
"""
import numpy as np
import pytest

import oracle as O
from pyevp import lib, synth
import ls_cases

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


# ---------------------------------------------------------------- FFT ----
@pytest.mark.parametrize("dims", [(8, 8, 8), (16, 16, 16), (32, 16, 8), (8, 32, 64), (64, 64, 64),
                                  (128, 128, 128), (256, 32, 16)])
def test_fft_r2c_fp64_vs_numpy(dims):
    x = np.random.default_rng(1).standard_normal(dims)
    ctx = lib.Context(*dims)
    assert rel(ctx.rfftn(x), np.fft.rfftn(x)) < 1e-12


@pytest.mark.parametrize("dims", [(8, 8, 8), (32, 16, 8), (128, 128, 128)])
def test_fft_c2r_fp64_vs_numpy(dims):
    rng = np.random.default_rng(2)
    nh = dims[2] // 2 + 1
    # arbitrary half spectrum, including imaginary parts on the kz = 0 and nz/2 planes
    X = rng.standard_normal(dims[:2] + (nh,)) + 1j * rng.standard_normal(dims[:2] + (nh,))
    ctx = lib.Context(*dims)
    n = np.prod(dims)
    ref = np.fft.irfftn(X, s=dims, axes=(0, 1, 2)) * n
    assert rel(ctx.irfftn_unnormalised(X), ref) < 1e-12


@pytest.mark.parametrize("dims", [(16, 16, 16), (128, 128, 128)])
def test_fft_fp32_vs_numpy(dims):
    x = np.random.default_rng(3).standard_normal(dims)
    ctx = lib.Context(*dims, precision=lib.MIXED)
    assert rel(ctx.rfftn(x), np.fft.rfftn(x)) < 2e-6
    X = np.fft.rfftn(x)
    assert rel(ctx.irfftn_unnormalised(X), x * np.prod(dims)) < 2e-6


# -------------------------------------------------------------- Green ----
def _loading_from(job):
    c066, s066, c0, s0 = job.setup_info()
    ld = lib.Loading()
    ld.c066[:] = c066.ravel()
    ld.s066[:] = s066.ravel()
    ld.c0[:] = c0.ravel()
    ld.s0[:] = s0.ravel()
    ld.dt = 1e-3
    return ld


@pytest.mark.parametrize("elast_bc", [1, 0])
@pytest.mark.parametrize("n", [8, 16])
def test_green_vs_oracle(job_factory, n, elast_bc):
    """GPU Gamma_eff on the half spectrum == Hermitian part of the oracle's
    Gamma on the full spectrum (what Re(IFFT) keeps)."""
    dims = (n, n, n)
    inp = job_factory(f"green{n}_{elast_bc}", n=n, grains=4, seed=11, increments=1,
                      elast_bc=elast_bc)
    J = O.Job(inp)
    rng = np.random.default_rng(5)
    s6 = rng.standard_normal((6,) + dims)
    full = np.empty(dims + (3, 3))
    idx = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    for c, (a, b) in enumerate(idx):
        full[..., a, b] = full[..., b, a] = s6[c]
    F = np.fft.fftn(full, axes=(0, 1, 2))
    orr, oi = J.green(F.real.reshape(-1, 3, 3), F.imag.reshape(-1, 3, 3))
    X = (orr + 1j * oi).reshape(dims + (3, 3))
    # Hermitian part H(k) = (X(k) + conj(X(-k)))/2
    m = [(-np.arange(d)) % d for d in dims]
    Xm = np.conj(X[m[0]][:, m[1]][:, :, m[2]])
    H = 0.5 * (X + Xm)
    nh = n // 2 + 1
    ctx = lib.Context(*dims)
    ld = _loading_from(J)
    ld.elast_bc = elast_bc
    ctx.set_loading(ld)
    Sh = np.stack([np.fft.rfftn(s6[c]) for c in range(6)])
    E = ctx.green(Sh)
    for a in range(3):
        for b in range(3):
            assert rel(E[a * 3 + b], H[:, :, :nh, a, b]) < 1e-12, (a, b)


# ------------------------------------------------------------- Newton ----
def _newton_batch(J, gj, rng, perturb):
    """Per-voxel inputs from an oracle state, with the strain perturbed so the
    Newton solve has work to do."""
    n = J.n
    sig = J.field(O.SIG)
    dg = J.field(O.DISGRAD)
    eps = J.field(O.EPS)
    tran = J.field(O.TRANMAT)
    crss = J.field(O.CRSS)
    ph = J.phase()
    cgr = J.cgr()
    c066, _, _, _ = J.setup_info()
    strain = 0.5 * (dg + dg.transpose(0, 2, 1))
    strain = strain + perturb * rng.standard_normal((n, 3, 3))
    strain = 0.5 * (strain + strain.transpose(0, 2, 1))
    return sig, strain, eps, tran, crss, ph, cgr, c066


def test_newton_vs_oracle(job_factory):
    inp = job_factory("newton8", n=8, grains=6, seed=21, increments=3, pore_fraction=0.05,
                      pore_radius=1)
    J = O.Job(inp)
    J.run(3)
    gj = lib.GpuJob(inp)
    rng = np.random.default_rng(7)
    sig, strain, eps, tran, crss, ph, cgr, c066 = _newton_batch(J, gj, rng, 2e-4)
    xt_schm = None
    # crystal voxels only
    sel = np.where(ph == 1)[0]
    m = len(sel)
    # Schmid tensors in sample axes: rotate the crystal-axis tensors with the oracle
    import oracle as Om
    from pyevp import synth
    nb = np.array([s[0] for s in synth.FCC_SLIP], float)
    bb = np.array([s[1] for s in synth.FCC_SLIP], float)
    nb /= np.linalg.norm(nb, axis=1)[:, None]
    bb /= np.linalg.norm(bb, axis=1)[:, None]
    xt_schm = np.array([Om.t2_to_v5(0.5 * (np.outer(nb[s], bb[s]) + np.outer(bb[s], nb[s])))
                        for s in range(12)])
    schm = np.array([Om.rotate_schmid(tran[v], xt_schm) for v in sel])       # m,12,5
    sg = np.array([Om.lu_inv66(cgr[v]) for v in sel])
    mat = synth.Material()
    nsr = np.full(12, mat.nrs)
    g0 = np.full(12, mat.gam0)
    # oracle per voxel
    ref = [Om.update_stress_voxel(sg[i], c066, 1e-3, sig[v], eps[v], strain[v], schm[i], crss[v],
                                  nsr, g0) for i, v in enumerate(sel)]
    assert all(r[0] == 0 for r in ref)
    # GPU batch (component-major arrays)
    def sym6(t):
        return np.stack([t[:, 0, 0], t[:, 1, 1], t[:, 2, 2], t[:, 1, 2], t[:, 0, 2], t[:, 0, 1]])
    st, s_out, ed, gd, err, nit = lib.Context.newton(
        gj_ctx(gj), np.ones(m, np.int32), np.arange(m, dtype=np.int32), sg.reshape(m, 36),
        sym6(sig[sel]), sym6(strain[sel]), sym6(eps[sel]),
        tran[sel].reshape(m, 9).T.copy(), crss[sel].reshape(m, 24).T.copy())
    assert st == 0
    s_ref = np.array([sym6(r[1][None])[:, 0] for r in ref]).T
    e_ref = np.array([sym6(r[2][None])[:, 0] for r in ref]).T
    g_ref = np.array([r[3] for r in ref]).T
    assert np.array_equal(nit, np.array([r[6] for r in ref]))
    assert rel(s_out, s_ref) < 1e-12
    assert rel(ed, e_ref) < 1e-10
    assert rel(gd, g_ref) < 1e-10
    assert rel(err[:, 0], np.array([r[4] for r in ref])) < 1e-8


class _CtxView:
    def __init__(self, h):
        self.h = h


def gj_ctx(gj):
    return _CtxView(gj.ctx)


# --------------------------------------------------------- whole loop ----
def _compare_jobs(J, gj, k, tol=1e-10):
    it_o = J.run(k)
    it_g, _ = gj.run(k)
    assert list(it_g) == list(it_o), (it_g, it_o)
    s = gj.state()
    assert rel(s["Sig"], J.field(O.SIG)) < tol
    assert rel(s["Edot"], J.field(O.EDOT)) < tol
    assert rel(s["Eps"], J.field(O.EPS)) < tol
    assert rel(s["DisGrad"], J.field(O.DISGRAD)) < tol
    assert rel(s["TranMat"], J.field(O.TRANMAT)) < tol
    assert rel(s["crss"], J.field(O.CRSS)) < tol
    mo, mg = J.macro(), gj.macro()
    assert rel(mg["sig_avg"], mo["SigAvg"]) < tol
    for a, b in (("e_vm", "e_vm"), ("s_vm", "s_vm"), ("evmp", "evmp"), ("dvmp", "dvmp")):
        assert abs(mg[a] - mo[b]) <= tol * max(abs(mo[b]), 1e-300), a
    g, p = gj.ids()
    assert np.array_equal(g, J.grain()) and np.array_equal(p, J.phase())


def test_loop_parity_8(job_factory):
    inp = job_factory("loop8", n=8, grains=4, seed=7, increments=5)
    _compare_jobs(O.Job(inp), lib.GpuJob(inp), 5)


def test_loop_parity_16_pores(job_factory):
    inp = job_factory("loop16p", n=16, grains=8, seed=9, increments=3, pore_fraction=0.03,
                      pore_radius=2)
    _compare_jobs(O.Job(inp), lib.GpuJob(inp), 3)


def test_loop_parity_full_size_128(tmp_path):
    """The 128^3 job of the benchmark (512 grains, 2 % pores) at full size:
    increment 1 forced to 3 Green iterations in the oracle (16 threads, a few
    seconds) and on the GPU, FP64 to the loop bar and the mixed loop to
    1e-6."""
    inp = synth.write_job(str(tmp_path), synth.config_c2())
    J = O.Job(inp)
    J.run(1, forced=[3])
    ref = J.field(O.SIG)
    # the strain rate goes as sigma^n (n = 10): its relative error bar in the
    # mixed loop is n times the stress's
    for prec, tol, tol_edot in ((lib.FP64, 1e-10, 1e-10), (lib.MIXED, 1e-6, 1e-5)):
        g = lib.GpuJob(inp, prec)
        g.run(1, forced=3)
        s = g.state()
        assert rel(s["Sig"], ref) < tol, (prec, rel(s["Sig"], ref))
        assert rel(s["Edot"], J.field(O.EDOT)) < tol_edot, (prec, rel(s["Edot"], J.field(O.EDOT)))
        assert rel(s["DisGrad"], J.field(O.DISGRAD)) < tol, (prec, rel(s["DisGrad"], J.field(O.DISGRAD)))
        g.close()


def test_loop_parity_simple_shear(job_factory):
    """Fully strain-controlled simple shear (no stress-controlled component,
    lattice spin from the 12 shear rate) against the oracle.  (ElastBC = 0 is
    checked on the Green operator itself: on this loading it drives the whole
    stress field to rounding level, where a loop comparison means nothing.)"""
    inp = job_factory("loopshear", n=16, grains=8, seed=43, increments=3, pore_fraction=0.02,
                      pore_radius=1, velgrad_bc_flag=(1, 1, 1, 1, 1, 1),
                      velgrad_bc=(0.0, 0.0, 0.0, 0.0, 0.0, 1.0), stress_bc_flag=(0, 0, 0, 0, 0, 0))
    _compare_jobs(O.Job(inp), lib.GpuJob(inp), 3)


@pytest.mark.parametrize("dims", [(16, 8, 8), (8, 16, 32)])
def test_loop_parity_non_cubic(job_factory, dims):
    """Non-cubic grids: nx != ny = nz runs the fused transforms, ny != nz the
    seven passes; both against the oracle, pores included."""
    inp = job_factory(f"loopnc{dims}", n=0, dims=dims, grains=6, seed=37, increments=3,
                      pore_fraction=0.02, pore_radius=1)
    _compare_jobs(O.Job(inp), lib.GpuJob(inp), 3)


def test_mixed_precision_non_cubic(job_factory):
    """Mixed loop on a 32 x 16 x 16 grid against the FP64 loop at equal counts."""
    inp = job_factory("mixnc", n=0, dims=(32, 16, 16), grains=8, seed=41, increments=2)
    a = lib.GpuJob(inp, lib.FP64)
    ia, _ = a.run(2)
    b = lib.GpuJob(inp, lib.MIXED)
    for k in range(2):
        b.run(1, forced=int(ia[k]))
    assert rel(b.state()["Sig"], a.state()["Sig"]) < 1e-6


@pytest.mark.parametrize("nrs2", [20.0, 12.5, 34.0])
def test_loop_parity_two_crystal_phases(job_factory, nrs2):
    """Two crystalline phases on alternating grains: waves that straddle a
    phase boundary run the Newton update once per phase.  The second phase's
    exponent selects the rate-law variant: n = 20 the fast one (integral
    n <= 32, one rate per phase), n = 12.5 pow(), n = 34 the squaring loop
    (on this job the reference's line search gives up, slope >= 0, from
    n = 36 on)."""
    m2 = synth.Material(c11=246500.0, c12=147300.0, c44=124700.0, nrs=nrs2, gam0=0.5,
                        tau0f=30.0, tau0b=30.0, tau1=15.0, thet0=300.0, thet1=30.0)
    inp = job_factory(f"loop2ph{nrs2}", n=16, grains=8, seed=29, increments=3, phase2=m2)
    _compare_jobs(O.Job(inp), lib.GpuJob(inp), 3)


@pytest.mark.parametrize("n,prec", [(16, lib.FP64), (32, lib.FP64), (32, lib.MIXED), (24, lib.FP64),
                                    (48, lib.MIXED)])
def test_fused_transforms_match_seven_passes(job_factory, n, prec):
    """The three fused FFT/Green passes against the seven-pass path (the one
    the FFT tests pin to numpy): same Green iteration counts, states equal to
    rounding."""
    inp = job_factory(f"fused{n}", n=n, grains=8, seed=19, increments=2, pore_fraction=0.02 if n == 32 else 0.0)
    a = lib.GpuJob(inp, prec)
    b = lib.GpuJob(inp, prec)
    b.set_fused_transforms(False)
    ia, _ = a.run(2)
    ib, _ = b.run(2)
    assert list(ia) == list(ib)
    A, B = a.state(), b.state()
    tol = 1e-10 if prec == lib.FP64 else 1e-6
    for k in ("Sig", "Edot", "DisGrad", "TranMat", "crss"):
        assert rel(A[k], B[k]) < tol, (k, rel(A[k], B[k]))
    assert rel(a.macro()["sig_avg"], b.macro()["sig_avg"]) < tol


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("prec", [lib.FP64, lib.MIXED])
def test_kernel_timing_groups(job_factory, fused, prec):
    """With HIP-event timing on, every kernel group of the path in use is
    counted once per Green iteration and no timer pair left unrecorded is
    queried (that would leave a HIP error for the next launch check)."""
    inp = job_factory("timing16", n=16, grains=8, seed=23, increments=1)
    job = lib.GpuJob(inp, prec)
    job.set_fused_transforms(fused)
    job.set_timing(True)
    it, _ = job.run(1)
    t = [job.timing(k) for k in range(6)]
    for k in (1, 2, 3):
        assert t[k][1] == t[0][1] - 1, (k, t)   # Newton also replays its last launch
        assert t[k][0] > 0.0
    assert t[0][1] == it[0] + 1
    assert t[4][1] == 2
    if prec == lib.MIXED:
        assert t[5][1] >= 1
    else:
        assert t[5][1] == 0


def test_kernel_timing_sample(job_factory):
    """The Newton timer alone, on one Green iteration in 4: about a quarter of
    the launches are timed (the increment's closing replay always), no other
    timer records, and the sampled average is a launch duration."""
    inp = job_factory("timing16", n=16, grains=8, seed=23, increments=1)
    job = lib.GpuJob(inp, lib.MIXED)
    job.set_timing(True, groups=[0], every=4)
    it, _ = job.run(1)
    ms, n = job.timing(0)
    assert abs((n - 1) - it[0] / 4) <= 2, (n, it)
    assert 0.0 < ms / n < 50.0
    for k in range(1, 6):
        assert job.timing(k)[1] == 0
    with pytest.raises(RuntimeError):
        job.set_timing(True, groups=[0], every=0)


def test_newton_shortcut_is_exact(job_factory):
    """Ending certainly-rejected Newton iterations early changes nothing: the
    whole state after 3 increments is bit-identical with the shortcut off."""
    inp = job_factory("short16", n=16, grains=8, seed=17, increments=3, pore_fraction=0.02)
    a = lib.GpuJob(inp)
    b = lib.GpuJob(inp)
    b.set_newton_shortcut(False)
    ia, sa = a.run(3)
    ib, sb = b.run(3)
    assert list(ia) == list(ib)
    assert [x["newton_iters"] for x in sa] == [x["newton_iters"] for x in sb]
    assert [x["ls_trials"] for x in sa] == [x["ls_trials"] for x in sb]
    A, B = a.state(), b.state()
    for k in A:
        assert np.array_equal(A[k], B[k]), k


@pytest.mark.parametrize("prec", [lib.FP64, lib.MIXED])
def test_loop_batches_are_exact(job_factory, prec):
    """The loop control runs on the device, so the host may queue iterations
    ahead of it: batches of 1, 2 or 4 iterations (two in flight, the ones
    queued past the loop's end doing nothing) give the state, the iteration
    counts and the Newton/line-search sums bit for bit of a host that waits
    for every iteration (batch 0), in converged and forced increments."""
    inp = job_factory("batch16", n=16, grains=8, seed=29, increments=4, pore_fraction=0.02)
    runs = []
    for batch in (0, 1, 2, 4):
        job = lib.GpuJob(inp, prec)
        job.set_loop_batch(batch)
        it, st = job.run(3)
        itf, stf = job.run(1, forced=5)
        runs.append((list(it) + list(itf), [x["newton_iters"] for x in st + stf],
                     [x["ls_trials"] for x in st + stf], [x["fp64_iters"] for x in st + stf],
                     job.state(), job.macro()))
    it0, nit0, ntr0, f0, s0, m0 = runs[0]
    assert it0[-1] == 5
    for it, nit, ntr, f, s, m in runs[1:]:
        assert it == it0 and nit == nit0 and ntr == ntr0 and f == f0
        for k in s0:
            assert np.array_equal(s[k], s0[k]), k
        for k in ("sig_avg", "s_vm", "e_vm", "err_s", "err_e"):
            assert np.array_equal(np.asarray(m[k]), np.asarray(m0[k])), k
    with pytest.raises(RuntimeError):
        lib.GpuJob(inp, prec).set_loop_batch(5)


def test_mixed_precision_stress_error(job_factory):
    """FP32 FFT/Green with FP64 Newton, anchors, projection and FP64 sweeps
    against the FP64 loop.  At equal iteration counts (the parity protocol
    for convergence tests that can flip) the difference is the precision
    error alone: max |sig_mix - sig_64| / max |sig_64| <= 1e-6.  Converging
    on their own, the two runs may stop one iteration apart, which moves the
    stress by about the loop tolerance (1e-6 here) itself."""
    inp = job_factory("mixed16", n=16, grains=8, seed=13, increments=3)
    a = lib.GpuJob(inp, lib.FP64)
    it, _ = a.run(3)
    b = lib.GpuJob(inp, lib.MIXED)
    for k in range(3):
        _, st = b.run(1, forced=int(it[k]))
        assert st[0]["fp64_iters"] >= 1
    sa, sb = a.state(), b.state()
    for f in ("Sig", "Edot", "DisGrad"):
        err = np.abs(sb[f] - sa[f]).max() / np.abs(sa[f]).max()
        print(f, err)
        assert err < 1e-6, (f, err)
    assert rel(b.macro()["sig_avg"], a.macro()["sig_avg"]) < 1e-8
    c = lib.GpuJob(inp, lib.MIXED)
    itc, stats = c.run(3)
    assert all(s["fp64_iters"] >= 1 for s in stats)
    assert np.abs(np.asarray(itc) - np.asarray(it)).max() <= 2
    assert rel(c.macro()["sig_avg"], a.macro()["sig_avg"]) < 1e-6


# ------------------------------------------------------ golden fixtures ----
def _golden_compare(inp, fx, nsteps, tol=1e-10):
    import os
    g = np.load(fx)
    gj = lib.GpuJob(inp)
    iters, sig_avg, vm = [], [], []
    for _ in range(nsteps):
        it, _ = gj.run(1)
        iters.append(int(it[0]))
        m = gj.macro()
        sig_avg.append(m["sig_avg"].ravel())
        vm.append([m[k] for k in ("time_tot", "e_vm", "evmp", "d_vm", "dvmp", "s_vm", "s_vm1")])
    assert iters == list(g["iters"]), (iters, list(g["iters"]))
    assert rel(np.array(sig_avg), g["sig_avg"]) < tol
    assert rel(np.array(vm), g["vm"]) < tol
    s = gj.state()
    sel = g["sel"]
    assert rel(s["Sig"][sel], g["sig"]) < tol
    assert rel(s["Edot"][sel], g["edot"]) < tol
    assert rel(s["DisGrad"][sel], g["disgrad"]) < tol
    assert rel(s["crss"][sel], g["crss"]) < tol
    assert rel(s["TranMat"][sel], g["tranmat"]) < tol
    gr, ph = gj.ids()
    assert np.array_equal(gr, g["grain"]) and np.array_equal(ph, g["phase"])


def test_golden_reference_test_run():
    """The reference's own test_run inputs (32^3, 24 % pores), both increments."""
    import os
    from conftest import GOLDEN
    _golden_compare(os.path.join(GOLDEN, "test_run", "input.in"),
                    os.path.join(GOLDEN, "oracle_test_run.npz"), 2)


def test_golden_c1_32(tmp_path):
    """32^3, 64 grains (the CPU-runnable configuration), all 10 increments."""
    import os
    from conftest import GOLDEN
    from pyevp import synth
    inp = synth.write_job(str(tmp_path), synth.config_c1())
    _golden_compare(inp, os.path.join(GOLDEN, "oracle_c1_32.npz"), synth.config_c1().increments)


def test_golden_16_pores_20_increments(tmp_path):
    """16^3 with 2 % pores, 20 increments: 18 orientation updates and
    hardening steps compounding, against the oracle's fixture."""
    import os
    from conftest import GOLDEN
    from pyevp import synth
    cfg = synth.Config(n=16, grains=8, seed=0x5EED0016, increments=20, pore_fraction=0.02, pore_radius=1)
    inp = synth.write_job(str(tmp_path), cfg)
    _golden_compare(inp, os.path.join(GOLDEN, "oracle_16_pores_20.npz"), 20)


def test_forge_dropin_writes_vm_out(tmp_path):
    """`forge input.in` end to end: the namelist echo and the vm.out rows."""
    import os
    import shutil
    import subprocess
    from conftest import GOLDEN
    for f in ("input.in", "Cu_p.sx", "FFT_microstructure_pore_197.txt"):
        shutil.copy(os.path.join(GOLDEN, "test_run", f), tmp_path / f)
    r = subprocess.run([os.path.join(lib.LIBDIR, "forge"), "input.in"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "NameList for input data" in r.stdout
    rows = open(tmp_path / "vm.out").read().splitlines()
    assert rows[0] == "TIME\tEVM\tEVMP\tDVM\tDVMP\tSVM\tSVMSVM_1"
    vals = np.array([[float(x) for x in ln.split("\t")] for ln in rows[1:]])
    g = np.load(os.path.join(GOLDEN, "oracle_test_run.npz"))
    assert vals.shape == (2, 7)
    # vm.out carries %e (7 significant digits)
    assert np.allclose(vals, g["vm"], rtol=1e-6, atol=0)


def test_forge_dumps_fields_on_the_reference_schedule(tmp_path):
    """PrintControl 1 in test_run/input.in: after increment 0 (0 % 500 == 0)
    the live loop dumps the elastic strain and the stress
    (src/evolution.c:1633-1648) in the .iout layout of src/io.c:283-476."""
    import os
    import shutil
    import subprocess
    from conftest import GOLDEN
    for f in ("input.in", "Cu_p.sx", "FFT_microstructure_pore_197.txt"):
        shutil.copy(os.path.join(GOLDEN, "test_run", f), tmp_path / f)
    r = subprocess.run([os.path.join(lib.LIBDIR, "forge"), "input.in"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    n = 32 ** 3
    for pre in ("els", "sig_afterDRX"):
        for c in range(1, 7):
            f = tmp_path / f"{pre}_{c}_S000000.iout"
            assert f.stat().st_size == 8 * n, f
        assert not (tmp_path / f"{pre}_1_S000001.iout").exists()
    # the stress dump after increment 0 equals the state then
    job = lib.GpuJob(str(tmp_path / "input.in"))
    job.run(1)
    s = job.state()["Sig"]
    ij = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    for c, (i, j) in enumerate(ij):
        d = np.fromfile(tmp_path / f"sig_afterDRX_{c + 1}_S000000.iout", dtype="<f8")
        assert np.array_equal(d, s[:, i, j])


def test_job_field_and_texture_dumps(job_factory, tmp_path):
    import ctypes as C
    inp = job_factory("dump8", n=8, grains=4, seed=21, increments=3)
    job = lib.GpuJob(inp)
    job.run(3)
    st = job.state()
    L = lib.evp()
    L.evp_job_write_field.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_int]
    L.evp_job_write_texture.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
    ij = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    dg = st["DisGrad"]
    els = 0.5 * (dg + dg.transpose(0, 2, 1)) - st["Eps"]
    expect = {0: st["Sig"], 1: st["Eps"], 2: st["Edot"], 3: els, 4: dg}
    for which, ref in expect.items():
        pre = str(tmp_path / f"f{which}").encode()
        assert L.evp_job_write_field(job.h, which, pre, 12) == 0
        for c, (i, j) in enumerate(ij):
            d = np.fromfile(tmp_path / f"f{which}_{c + 1}_S000012.iout", dtype="<f8")
            assert np.abs(d - ref[:, i, j]).max() <= 1e-15 * max(np.abs(ref).max(), 1.0), (which, c)
    assert L.evp_job_write_texture(job.h, str(tmp_path / "tex").encode(), 3) == 0
    tex = np.dtype([("angle", "<f8", (3,)), ("coord", "<i4", (3,)), ("grain", "<i4"),
                    ("phase", "<i4"), ("buffer", "<i4")])
    rec = np.fromfile(tmp_path / "tex_S000003.iout", dtype=tex)
    g, p = job.ids()
    assert np.array_equal(rec["grain"], g) and np.array_equal(rec["phase"], p)
    # the angles rebuild each voxel's rotation (away from the poles)
    import oracle as O
    for v in range(0, job.n, 37):
        a = O.euler_to_mat(*(rec["angle"][v] * 3.14159265359 / 180.0))
        if abs(a[2, 2]) < 0.999:
            assert np.abs(a.T - st["TranMat"][v]).max() < 1e-9


@pytest.mark.parametrize("nrs", ls_cases.NRS_VALUES)
def test_newton_line_search_branches_vs_oracle(tmp_path, nrs):
    """The voxels of the oracle's line-search pins (tests/ls_cases.py:
    orientations, stresses and strains over several decades, n = 10, 20, 30)
    through evpgpu_newton_host: the oracle's trace records which branches
    each solve takes (accepted steps, TOLX rejections, quadratic and cubic
    backtracks); the GPU must give the same Newton count and the same stress.
    Solves whose Jacobian is nearly singular at some iterate (condition
    > 1e6) or whose merit value exceeds 1e12 are left out: there the two
    solves (Cholesky of the symmetric form on the GPU, the explicit pivoted
    LU inverse in the oracle) round differently enough to change step
    decisions, and at rates of |x|^n ~ 1e30 and beyond the Cholesky form can
    lose positive definiteness to rounding (status 35) where the pivoted LU
    goes on -- states no loop reaches."""
    import test_oracle_line_search as TL
    cfg = synth.Config(n=8, grains=4, seed=5, increments=1, material=synth.Material(nrs=nrs))
    gj = lib.GpuJob(synth.write_job(str(tmp_path), cfg))
    c066 = gj.setup_info()[0]
    L = TL._lib()
    vox, refs = [], []
    for (T, lam, E, crss) in ls_cases.cases(200, 500 + int(nrs), nrs):
        vx = TL.Voxel(T, lam, E, crss, nrs, c066)
        st, s, ed, gd, nit, rows = TL.solve(vx, L)
        if st:
            continue
        good = bool(np.all(rows[:, 3] < 1e12))
        if good:
            for r in rows[rows[:, 0] != np.roll(rows[:, 0], 1)]:
                if np.linalg.cond(vx.jac(r[8:14])) > 1e6:
                    good = False
                    break
        vox.append(vx)
        refs.append((s, nit, rows, good))
    n_all = len(vox)
    keep = [i for i in range(n_all) if refs[i][3]]
    vox = [vox[i] for i in keep]
    refs = [refs[i] for i in keep]
    m = len(vox)
    sym = lambda t: np.array([t[0, 0], t[1, 1], t[2, 2], t[1, 2], t[0, 2], t[0, 1]])
    sgr = np.array([O.lu_inv66(v.Cgr66) for v in vox]).reshape(m, 36)
    sig = np.array([sym(v.from6(v.lam6)) for v in vox]).T.copy()
    strain = np.array([sym(v.from6(v.E6)) for v in vox]).T.copy()
    tran = np.array([v.T.reshape(9) for v in vox]).T.copy()
    crss = np.array([v.crss.reshape(24) for v in vox]).T.copy()
    st, s_out, ed, gd, err, nit = lib.Context.newton(
        gj_ctx(gj), np.ones(m, np.int32), np.arange(m, dtype=np.int32), sgr, sig, strain,
        np.zeros((6, m)), tran, crss)
    assert st == 0
    seen = np.zeros(7, int)
    for i, (s, it, rows, good) in enumerate(refs):
        ref = sym(s)
        assert nit[i] == it, (i, nit[i], it)
        assert np.max(np.abs(s_out[:, i] - ref)) <= 1e-10 * np.max(np.abs(ref)), (i, s_out[:, i], ref)
        seen += np.bincount(rows[:, 6].astype(int), minlength=7)
    print(f"n = {nrs}: {m} of {n_all} voxels compared; branches of their solves {seen}")
    assert seen[0] > 0 and seen[1] > 0 and seen[2] > 0 and seen[5] + seen[6] > 0
    gj.close()


def test_plastic_laminate_vs_exact_solution(tmp_path):
    """The GPU loop on a two-crystal laminate along x (4 x 2 x 2 voxels) over
    10 increments against the exact laminate solution solved in numpy
    (tests/crystal_np.py; the oracle's own check in test_oracle_plastic.py):
    stress and displacement gradient of every increment, then plastic strain,
    orientation and CRSS."""
    import os
    import crystal_np as X
    mat = synth.Material(nrs=12.0, tau1=20.0, thet0=300.0)
    eulers = [(10.0, 30.0, 40.0), (70.0, 20.0, 110.0)]
    dims = (4, 2, 2)
    cfg = synth.Config(n=0, dims=dims, grains=2, increments=10, err=1e-12, iter_max=40000, material=mat)
    N = 16
    lay = (np.arange(N) // 4) < 2
    eul = np.where(lay[:, None], np.array(eulers[0]), np.array(eulers[1]))
    synth.write_ms(os.path.join(str(tmp_path), "ms.txt"), dims, eul, np.where(lay, 1, 2).astype(np.int32),
                   np.ones(N, np.int32))
    synth.write_sx(os.path.join(str(tmp_path), "Cu.sx"), mat)
    synth.write_input(os.path.join(str(tmp_path), "input.in"), cfg)
    gj = lib.GpuJob(os.path.join(str(tmp_path), "input.in"))
    gj.set_newton_tolerances(1e-14, 1e-13)
    M = X.Laminate(eulers, mat)
    for k in range(10):
        gj.run(1)
        sigs, res = M.increment()
        st = gj.state()
        ref = np.where(lay[:, None, None], sigs[0][None], sigs[1][None])
        assert np.max(np.abs(st["Sig"] - ref)) <= 1e-9 * np.max(np.abs(ref)), k
        grad = np.where(lay[:, None, None], M.L[0].grad[None], M.L[1].grad[None])
        assert np.max(np.abs(st["DisGrad"] - grad)) <= 1e-9 * np.max(np.abs(grad)), k
    for l, Ly in enumerate(M.L):
        v = 0 if l == 0 else N - 1
        assert np.max(np.abs(st["TranMat"][v] - Ly.T)) <= 1e-9
        assert np.max(np.abs(st["crss"][v] - Ly.crss[:, None])) <= 1e-9 * np.max(Ly.crss)
        assert np.max(np.abs(st["Eps"][v] - Ly.epsp)) <= 1e-9 * np.max(np.abs(Ly.epsp))
    gj.close()
