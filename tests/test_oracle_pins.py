"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no golden outputs and cannot be built here (it needs MPI,
GSL and FFTW-MPI headers the image lacks), so every oracle function is pinned
by a closed form, an invariant, a library routine or an independent
formulation, each chosen so that a plausible slip (a dropped factor, a wrong
sign or index, a transposed operand) fails.
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle as O
from pyevp import synth

RNG = np.random.default_rng(1234)


def cubic_c66(c11, c12, c44):
    return O.t4_to_c2(cubic_t4(c11, c12, c44))


def cubic_t4(c11, c12, c44):
    c = np.zeros((3, 3, 3, 3))
    for i in range(3):
        c[i, i, i, i] = c11
        for j in range(3):
            if i != j:
                c[i, i, j, j] = c12
                c[i, j, i, j] = c[i, j, j, i] = c44
    return c


# ------------------------------------------------------------ basis -----
def test_basis_is_orthonormal_and_matches_lebensohn_tome():
    B = O.basis()
    G = np.einsum("iab,jab->ij", B, B)
    assert np.allclose(G, np.eye(6), atol=1e-15)
    assert np.allclose(B, B.transpose(0, 2, 1))
    assert np.allclose(B[5], np.eye(3) / math.sqrt(3), atol=1e-16)
    assert np.allclose(np.trace(B[:5], axis1=1, axis2=2), 0, atol=1e-16)
    assert np.isclose(B[0, 1, 1], 1 / math.sqrt(2)) and np.isclose(B[1, 2, 2], 2 / math.sqrt(6))


def test_basis_roundtrips():
    t = RNG.standard_normal((3, 3))
    t = t + t.T
    assert np.allclose(O.v6_to_t2(O.t2_to_v6(t)), t, atol=1e-14)
    d = t - np.trace(t) / 3 * np.eye(3)
    assert np.allclose(O.v5_to_t2(O.t2_to_v5(d)), d, atol=1e-14)
    c = RNG.standard_normal((6, 6))
    assert np.allclose(O.t4_to_c2(O.c2_to_t4(c)), c, atol=1e-13)
    # the 6x6 form contracts like the tensor: (C:e) in the basis == C66 e6
    c = c + c.T
    e = RNG.standard_normal((3, 3))
    e = e + e.T
    lhs = O.t2_to_v6(np.einsum("abcd,cd->ab", O.c2_to_t4(c), e))
    assert np.allclose(lhs, c @ O.t2_to_v6(e), atol=1e-12)


# -------------------------------------------------------- rotations -----
def test_euler_matrix_is_bunge_rotation():
    assert np.allclose(O.euler_to_mat(0, 0, 0), np.eye(3))
    for _ in range(5):
        a = RNG.uniform(0, 2 * math.pi, 3)
        R = O.euler_to_mat(*a)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-15)
        assert np.isclose(np.linalg.det(R), 1.0)
        # Bunge (z-x-z) passive rotation: sample -> crystal
        ref = Rotation.from_euler("ZXZ", a).as_matrix().T
        assert np.allclose(R, ref, atol=1e-14)


def test_ten4_transform_invariances():
    iso = cubic_t4(200.0, 80.0, 60.0)  # c44 = (c11 - c12)/2: isotropic
    R = Rotation.random(random_state=3).as_matrix()
    assert np.allclose(O.ten4_transform_inv(R, iso), iso, atol=1e-11)
    cub = cubic_t4(168.4, 121.4, 75.4)
    R90 = Rotation.from_euler("z", 90, degrees=True).as_matrix()
    assert np.allclose(O.ten4_transform_inv(R90, cub), cub, atol=1e-12)
    # c1_ijkl = a_mi a_nj a_pk a_ql c2_mnpq, written with einsum
    ref = np.einsum("mi,nj,pk,ql,mnpq->ijkl", R, R, R, R, cub)
    assert np.allclose(O.ten4_transform_inv(R, cub), ref, atol=1e-11)


def test_lu_inverse_and_its_quirks():
    a = RNG.standard_normal((6, 6)) + 6 * np.eye(6)
    assert np.allclose(O.lu_inv66(a), np.linalg.inv(a), atol=1e-14)
    # row maximum below 1e-5 -> the reference exits with 35 (src/kinematics.c:306-313)
    s = np.eye(6)
    s[3] = 1e-6
    with pytest.raises(O.OracleError) as e:
        O.lu_inv66(s)
    assert e.value.code == 35
    # a pivot below 1e-5 is replaced by 1e-20 (src/kinematics.c:345): the inverse of this
    # nearly singular block is ~1e20, not the 1e7 numpy returns
    q = np.eye(6)
    q[0, 0] = q[0, 1] = q[1, 0] = 1.0
    q[1, 1] = 1.0 + 1e-7
    inv = O.lu_inv66(q)
    assert np.max(np.abs(inv)) > 1e19
    assert np.max(np.abs(np.linalg.inv(q))) < 1e8


def test_vonmises_closed_forms():
    assert np.isclose(O.vonmises(np.diag([1.0, -0.5, -0.5])), 1.0)
    g = np.zeros((3, 3))
    g[0, 1] = 1.0  # simple shear: sqrt(2/3 * 2 * (1/2)^2) = 1/sqrt(3)
    assert np.isclose(O.vonmises(g), 1 / math.sqrt(3))
    assert np.isclose(O.vonmises(np.eye(3) * 5.0), 0.0)


def test_orientation_update_is_axis_angle_rotation():
    T = Rotation.random(random_state=5).as_matrix()
    w = np.array([0.03, -0.02, 0.05])
    W = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    tnew, rot = O.orientation(T, W)
    assert np.allclose(rot, Rotation.from_rotvec(w).as_matrix(), atol=1e-14)
    assert np.allclose(tnew, rot @ T, atol=1e-14)
    # below |w| = 1e-6 the reference keeps |w| = 1 as the divisor
    # (src/constitutive.c:157-161): the applied spin is tan(|w|/2) * w, not w
    w0 = 2e-7
    W = np.array([[0, 0, 0], [0, 0, -w0], [0, w0, 0]])
    _, rot = O.orientation(np.eye(3), W)
    t = math.tan(w0 / 2)
    assert math.isclose(rot[2, 1], 2 * (w0 * t) / (1 + t * t), rel_tol=1e-12)
    assert rot[2, 1] < 1e-6 * math.sin(w0)


# ------------------------------------------------------ constitutive -----
def fcc_schmid_xt():
    nb = np.array([s[0] for s in synth.FCC_SLIP], float)
    bb = np.array([s[1] for s in synth.FCC_SLIP], float)
    nb /= np.linalg.norm(nb, axis=1)[:, None]
    bb /= np.linalg.norm(bb, axis=1)[:, None]
    return np.array([O.t2_to_v5(0.5 * (np.outer(nb[s], bb[s]) + np.outer(bb[s], nb[s])))
                     for s in range(12)])


def test_strain_rate_schmid_closed_form_001():
    """[001] tension: 8 of the 12 {111}<110> systems have Schmid factor m = 1/sqrt(6),
    the other 4 zero.  At sigma_zz = sqrt(6) tau_c each active system runs at
    gamma0, so edot_zz = sum m gamma = 8 gamma0/sqrt(6) and edot_xx = edot_yy = -4/sqrt(6)."""
    sc = fcc_schmid_xt()
    tau_c = 16.0
    s = np.zeros((3, 3))
    s[2, 2] = math.sqrt(6) * tau_c
    ed, dd, gd = O.strain_rate(O.t2_to_v6(s), sc, np.full((12, 2), tau_c), np.full(12, 10.0),
                               np.full(12, 1.0))
    assert np.allclose(np.sort(np.abs(gd)), [0] * 4 + [1.0] * 8, atol=1e-12)
    E = O.v6_to_t2(ed)
    assert np.isclose(E[2, 2], 8.0 / math.sqrt(6), rtol=1e-12)
    assert np.isclose(E[0, 0], -4.0 / math.sqrt(6), rtol=1e-12)
    assert np.isclose(E[1, 1], -4.0 / math.sqrt(6), rtol=1e-12)
    assert np.max(np.abs(E - np.diag(np.diag(E)))) < 1e-14


def test_strain_rate_derivative_matches_finite_differences():
    sc = np.array([O.rotate_schmid(Rotation.random(random_state=9).as_matrix(), fcc_schmid_xt())])[0]
    crss = np.column_stack([np.full(12, 16.0), np.full(12, 18.0)])
    s6 = RNG.standard_normal(6) * 30
    ed, dd, _ = O.strain_rate(s6, sc, crss, np.full(12, 10.0), np.full(12, 1.0))
    h = 1e-5
    fd = np.zeros((6, 6))
    for j in range(6):
        e = np.zeros(6)
        e[j] = h
        fd[:, j] = (O.strain_rate(s6 + e, sc, crss, np.full(12, 10.0), np.full(12, 1.0))[0] -
                    O.strain_rate(s6 - e, sc, crss, np.full(12, 10.0), np.full(12, 1.0))[0]) / (2 * h)
    assert np.allclose(dd, fd, rtol=1e-6, atol=1e-9 * np.max(np.abs(dd)))
    assert np.all(dd[5] == 0) and np.all(dd[:, 5] == 0) and ed[5] == 0


def test_voce_hardening_is_exact_integral():
    """With unit self/latent hardening, each system's CRSS increment equals
    tau(G + dG) - tau(G) of the Voce law tau = tau0 + (tau1 + th1 G)(1 - exp(-th0 G/tau1))."""
    tau0, tau1, th0, th1 = 16.0, 10.0, 200.0, 20.0
    voce = lambda G: tau0 + (tau1 + th1 * G) * (1 - math.exp(-th0 * G / tau1))
    gd = RNG.uniform(-1, 1, 12)
    dt = 1e-3
    G0 = 0.013
    crss0 = np.full((12, 2), voce(G0))
    crss, g1 = O.harden_voxel(gd, dt, np.ones((12, 12)), np.tile([tau0, tau0, tau1], (12, 1)),
                              np.tile([th0, th1], (12, 1)), crss0, G0)
    dG = np.sum(np.abs(gd)) * dt
    assert np.isclose(g1, G0 + dG, rtol=1e-15)
    assert np.allclose(crss, voce(G0 + dG), rtol=1e-12)
    # zero slip leaves the CRSS unchanged (the reference computes 0/0 there)
    crss2, g2 = O.harden_voxel(np.zeros(12), dt, np.ones((12, 12)),
                               np.tile([tau0, tau0, tau1], (12, 1)), np.tile([th0, th1], (12, 1)),
                               crss0, G0)
    assert np.all(crss2 == crss0) and g2 == G0


# ---------------------------------------------------- stress update -----
def _voxel_inputs(elastic=False):
    R = Rotation.random(random_state=17).as_matrix()
    c66 = cubic_c66(168400., 121400., 75400.)
    cgr = O.t4_to_c2(O.ten4_transform_inv(R, O.c2_to_t4(c66)))
    sg = O.lu_inv66(cgr)
    c066 = cubic_c66(168400., 121400., 75400.) * 0.9 + np.diag([1000.] * 6)
    sc = O.rotate_schmid(R.T, fcc_schmid_xt())
    tau = 1e9 if elastic else 16.0
    crss = np.full((12, 2), tau)
    sig = RNG.standard_normal((3, 3)) * 20
    sig = sig + sig.T
    eps = RNG.standard_normal((3, 3)) * 1e-4
    eps = eps + eps.T
    eps -= np.trace(eps) / 3 * np.eye(3)
    dg = RNG.standard_normal((3, 3)) * 1e-3
    return sg, c066, sig, eps, dg, sc, crss


def test_update_stress_elastic_is_linear_solve():
    """Elastic limit: R(s) = s - lam + C0 (eps_p + S s - e) = 0 is linear."""
    sg, c066, sig, eps, dg, sc, crss = _voxel_inputs(elastic=True)
    st, s_out, ed, gd, es, ee, nit = O.update_stress_voxel(sg, c066, 1e-3, sig, eps, dg, sc, crss,
                                                           np.full(12, 10.0), np.full(12, 1.0))
    assert st == 0
    lam6 = O.t2_to_v6(sig)
    e6 = O.t2_to_v6(0.5 * (dg + dg.T))
    p6 = O.t2_to_v6(eps)
    s6 = np.linalg.solve(np.eye(6) + c066 @ sg, lam6 + c066 @ (e6 - p6))
    assert np.allclose(O.t2_to_v6(s_out), s6, rtol=1e-12, atol=1e-12 * np.max(np.abs(s6)))
    # Newton steps are capped at length 10 (stpmax, src/evolution.c:182); on this
    # linear problem every step points at the root, so the count is
    # ceil(|s - lambda| / 10) steps plus the final rejected one (below TOLX)
    d = np.linalg.norm(s6 - lam6)
    assert d > 10.0
    assert nit == math.ceil(d / 10.0) + 1


def test_update_stress_converged_root_satisfies_residual():
    sg, c066, sig, eps, dg, sc, crss = _voxel_inputs()
    dt = 1e-3
    st, s_out, ed, gd, es, ee, nit = O.update_stress_voxel(sg, c066, dt, sig, eps, dg, sc, crss,
                                                           np.full(12, 10.0), np.full(12, 1.0))
    assert st == 0 and 1 <= nit < 50
    s6 = O.t2_to_v6(s_out)
    edot6, _, _ = O.strain_rate(s6, sc, crss, np.full(12, 10.0), np.full(12, 1.0))
    res = s6 - O.t2_to_v6(sig) + c066 @ (O.t2_to_v6(eps) + edot6 * dt + sg @ s6 -
                                         O.t2_to_v6(0.5 * (dg + dg.T)))
    # the line search stops once a Newton step is below TOLX = 1e-7 (relative)
    assert np.max(np.abs(res)) < 1e-6 * np.max(np.abs(s6))
    assert np.isclose(es, np.linalg.norm(s6 - O.t2_to_v6(sig)), rtol=1e-12)


# --------------------------------------------------------------- FFT -----
@pytest.mark.parametrize("dims", [(4, 4, 4), (8, 4, 2), (6, 5, 3), (16, 16, 16)])
def test_fft_matches_numpy(dims):
    x = RNG.standard_normal(dims) + 1j * RNG.standard_normal(dims)
    assert np.allclose(O.fft3d(x, -1), np.fft.fftn(x), atol=1e-12 * np.sqrt(x.size))
    assert np.allclose(O.fft3d(x, +1), np.fft.ifftn(x) * x.size, atol=1e-12 * np.sqrt(x.size))


# ------------------------------------------------------------- Green -----
def _iso_job(job_factory, elast_bc=1):
    m = synth.Material(c11=200000., c12=80000., c44=60000.)
    return O.Job(job_factory(f"iso{elast_bc}", n=8, grains=3, seed=3, increments=1, material=m,
                             elast_bc=elast_bc))


def test_green_isotropic_closed_form(job_factory):
    """Isotropic C0 (lambda, mu): Omega = (I - (l+mu)/(l+2mu) n n)/mu, Gamma = -Omega n n."""
    J = _iso_job(job_factory)
    lam, mu = 80000., 60000.
    n = 8
    k = np.fft.fftfreq(n) * n
    ks = np.stack(np.meshgrid(k, k, k, indexing="ij"), -1).reshape(-1, 3)
    xi = ks / n - np.floor(ks / n + 0.5)
    xi[np.isclose(xi, 0.5)] = -0.5
    nrm = np.linalg.norm(xi, axis=1)
    nrm[0] = 1
    nv = xi / nrm[:, None]
    s = RNG.standard_normal((n ** 3, 3, 3)) + 1j * RNG.standard_normal((n ** 3, 3, 3))
    orr, oi = J.green(s.real, s.imag)
    Om = (np.eye(3)[None] - (lam + mu) / (lam + 2 * mu) * nv[:, :, None] * nv[:, None, :]) / mu
    ref = -np.einsum("vik,vj,vl,vkl->vij", Om, nv, nv, s)
    ref[0] = 0
    assert np.allclose(orr + 1j * oi, ref, rtol=1e-11, atol=1e-11 * np.max(np.abs(ref)))


def test_green_operator_invariants(job_factory):
    """For any C0: Gamma:s = 0 when s.n = 0, and Gamma:(C0:sym(a x n)) = -a x n (k != 0)."""
    J = O.Job(job_factory("ginv", n=8, grains=5, seed=4, increments=1))
    _, _, C0, _ = J.setup_info()
    n = 8
    v = 77  # some k != 0
    x, y, z = v // 64, (v // 8) % 8, v % 8
    xi = np.array([x, y, z]) / n
    xi -= np.floor(xi + 0.5)
    nv = xi / np.linalg.norm(xi)
    a = RNG.standard_normal(3)
    sym = 0.5 * (np.outer(a, nv) + np.outer(nv, a))
    s = np.zeros((n ** 3, 3, 3), complex)
    s[v] = np.einsum("ijkl,kl->ij", C0, sym)
    orr, oi = J.green(s.real, s.imag)
    assert np.allclose(orr[v], -np.outer(a, nv), atol=1e-14)
    t = RNG.standard_normal((3, 3))
    P = np.eye(3) - np.outer(nv, nv)
    s[v] = P @ (t + t.T) @ P  # s.n = 0
    orr, oi = J.green(s.real, s.imag)
    assert np.max(np.abs(orr[v])) < 1e-16 * 1e3


# -------------------------------------------------------- whole loop -----
def test_uniform_single_crystal_fields_are_uniform(job_factory):
    J = O.Job(job_factory("sx", n=8, grains=3, seed=5, increments=3,
                          single_orientation=(20.0, 35.0, 50.0)))
    J.run(3)
    s = J.field(O.SIG)
    assert np.max(np.abs(s - s[0])) < 1e-9 * np.max(np.abs(s))
    m = J.macro()
    # free lateral stresses
    assert abs(m["SigAvg"][0, 0]) < 1e-5 * m["s_vm"] and abs(m["SigAvg"][1, 1]) < 1e-5 * m["s_vm"]


def test_elastic_laminate_exact_solution(job_factory, tmp_path):
    """Two crystals layered along x, elastic: the converged stress is the exact
    laminate solution (traction sigma.e_x and in-plane strain continuous)."""
    from pyevp.synth import write_input, write_sx, write_ms, Config, Material
    n = 8
    mat = Material(tau0f=1e9, tau0b=1e9)
    cfg = Config(n=n, grains=2, increments=1, material=mat, err=1e-12)
    e1, e2 = (10.0, 30.0, 40.0), (70.0, 20.0, 110.0)
    idx = np.arange(n ** 3)
    lay = (idx // (n * n)) < n // 2
    eul = np.where(lay[:, None], np.array(e1), np.array(e2))
    grain = np.where(lay, 1, 2).astype(np.int32)
    write_ms(str(tmp_path / "ms.txt"), n, eul, grain, np.ones(n ** 3, np.int32))
    write_sx(str(tmp_path / "Cu.sx"), mat)
    write_input(str(tmp_path / "input.in"), cfg)
    J = O.Job(str(tmp_path / "input.in"))
    J.run(1)
    s = J.field(O.SIG)
    # exact laminate: unknowns per layer e11, e12, e13 (3 x 2) and E11, E22 (macro);
    # eq: sigma_1j equal in both layers (3), <e_1j> = E_1j with E12=E13=0 (3), <s11>=<s22>=0 (2)
    cgr = J.cgr()
    B = O.basis()
    C = [np.einsum("ij,iab,jcd->abcd", cgr[0 if l else -1], B, B) for l in (True, False)]
    dt = 1e-3
    E33 = dt
    import itertools

    def stress(Cl, e):
        return np.einsum("abcd,cd->ab", Cl, e)

    def layer_strain(u, E11, E22):
        e11, e12, e13 = u
        return np.array([[e11, e12, e13], [e12, E22, 0.0], [e13, 0.0, E33]])

    from scipy.optimize import fsolve

    def F(p):
        u1, u2, E11, E22 = p[0:3], p[3:6], p[6], p[7]
        s1 = stress(C[0], layer_strain(u1, E11, E22))
        s2 = stress(C[1], layer_strain(u2, E11, E22))
        return np.concatenate([s1[0] - s2[0], 0.5 * (u1 + u2) - np.array([E11, 0, 0]),
                               [0.5 * (s1[0, 0] + s2[0, 0]), 0.5 * (s1[1, 1] + s2[1, 1])]])

    p = fsolve(F, np.zeros(8), xtol=1e-14)
    s1 = stress(C[0], layer_strain(p[0:3], p[6], p[7]))
    s2 = stress(C[1], layer_strain(p[3:6], p[6], p[7]))
    ref = np.where(lay[:, None, None], s1[None], s2[None])
    assert np.max(np.abs(s - ref)) < 1e-7 * np.max(np.abs(ref))


def test_elastic_loop_matches_independent_numpy_loop(job_factory):
    """The elastic Green loop (Newton reduces to a linear solve) restated with
    numpy's FFT and vectorised tensors reproduces the oracle's per-iteration
    macro stress history."""
    m = synth.Material(tau0f=1e9, tau0b=1e9)
    inp = job_factory("el8np", n=8, grains=4, seed=7, increments=1, material=m)
    n, N = 8, 512
    hist_o = []
    for k in (1, 5, 20):
        J = O.Job(inp)
        J.run(1, forced=[k])
        hist_o.append(J.macro()["SigAvg"].copy())
    J = O.Job(inp)
    c066, s066, C0, S0 = J.setup_info()
    Cgr = J.cgr()
    B = O.basis()
    S66 = np.linalg.inv(Cgr)
    dt = 1e-3
    Ud = np.zeros((3, 3))
    Ud[2, 2] = 1
    E = np.zeros((N, 3, 3)) + Ud * dt
    C4 = np.einsum("vij,iab,jcd->vabcd", Cgr, B, B)
    sig = np.einsum("vabcd,cd->vab", C4, Ud * dt)
    k = np.fft.fftfreq(n) * n
    xi = np.stack(np.meshgrid(k, k, k, indexing="ij"), -1).reshape(-1, 3) / n
    xi = xi - np.floor(xi + 0.5)
    xi[np.isclose(xi, 0.5)] = -0.5
    nrm = np.linalg.norm(xi, axis=1)
    nrm[0] = 1
    nv = xi / nrm[:, None]
    A = np.einsum("ijkl,vj,vl->vik", C0, nv, nv)
    A[0] = np.eye(3)
    G = -np.einsum("vik,vj,vl->vijkl", np.linalg.inv(A), nv, nv)
    G[0] = 0
    IJV = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    dE = np.zeros((3, 3))
    hist = {}
    for it in range(1, 21):
        lam = sig.copy()
        ks = np.fft.fftn(sig.reshape(n, n, n, 3, 3), axes=(0, 1, 2)).reshape(N, 3, 3)
        de = np.fft.ifftn(np.einsum("vijkl,vkl->vij", G, ks).reshape(n, n, n, 3, 3),
                          axes=(0, 1, 2)).real.reshape(N, 3, 3)
        E = E + dE + de
        e6 = np.einsum("vab,iab->vi", 0.5 * (E + E.transpose(0, 2, 1)), B)
        l6 = np.einsum("vab,iab->vi", lam, B)
        s6 = np.linalg.solve(np.eye(6)[None] + np.einsum("ij,vjk->vik", c066, S66),
                             (l6 + e6 @ c066.T)[..., None])[..., 0]
        sig = np.einsum("vi,iab->vab", s6, B)
        SA = sig.mean(0)
        dE = np.zeros((3, 3))
        for i, (a, b) in enumerate(IJV[:2]):
            for kk, (c, d) in enumerate(IJV[:2]):
                dE[a, b] += S0[a, b, c, d] * (0 - SA[c, d])
        hist[it] = SA
    for k, ref in zip((1, 5, 20), hist_o):
        assert np.allclose(hist[k], ref, rtol=1e-10, atol=1e-10 * np.max(np.abs(ref)))


def test_grain_relabelling_leaves_fields_unchanged(job_factory, tmp_path):
    inp = job_factory("perm", n=8, grains=4, seed=8, increments=2)
    J = O.Job(inp)
    J.run(2)
    import os
    import shutil
    d = os.path.dirname(inp)
    for f in ("input.in", "Cu.sx"):
        shutil.copy(os.path.join(d, f), tmp_path / f)
    lines = open(os.path.join(d, "ms.txt")).read().splitlines()
    perm = {1: 3, 2: 1, 3: 4, 4: 2}
    with open(tmp_path / "ms.txt", "w") as f:
        for ln in lines:
            p = ln.split()
            p[6] = str(perm.get(int(p[6]), int(p[6])))
            f.write(" ".join(p) + "\n")
    J2 = O.Job(str(tmp_path / "input.in"))
    J2.run(2)
    assert np.array_equal(J2.field(O.SIG), J.field(O.SIG))


# ---------------------------------------------------- reference files -----
def test_reference_fixture_setup():
    """test_run/: 32^3 single crystal (grain 2, phase 1, Euler 0) with 7968 gas
    voxels (grain 1, phase 2); C066 = (24800/32768) C in the basis (gas as zero)."""
    from conftest import GOLDEN
    import os
    J = O.Job(os.path.join(GOLDEN, "test_run", "input.in"))
    g, p = J.grain(), J.phase()
    assert np.sum(p == 2) == 7968 and np.all(g[p == 2] == 1) and np.all(g[p == 1] == 2)
    c066, s066, _, _ = J.setup_info()
    ref = cubic_c66(162700., 119000., 71150.) * (24800 / 32768)
    assert np.allclose(c066, ref, rtol=1e-14, atol=1e-9)
    assert np.allclose(s066 @ c066, np.eye(6), atol=1e-12)
    assert np.isclose(J.macro()["Wgt_ph1"], 24800 / 32768)
