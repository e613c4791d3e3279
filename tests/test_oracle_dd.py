"""Pins of the oracle's dislocation-density law (the law the reference's
shipped build runs: src/evpFlags.h DD_BASED_FLAG, DD_CU, FORWARD_PLUS_BACKWARD,
DD_GND_FIRST_ORDER).  Each check is fixed by geometry, calculus or a closed
form rather than by the formulas the oracle restates."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
sys.path.insert(0, os.path.join(HERE, "..", "m-themes_amd"))

import oracle as O  # noqa: E402
from pyevp import synth  # noqa: E402

FCC_N = np.array([s[0] for s in synth.FCC_SLIP], float)
FCC_B = np.array([s[1] for s in synth.FCC_SLIP], float)
N_HAT = FCC_N / np.linalg.norm(FCC_N, axis=1)[:, None]
B_HAT = FCC_B / np.linalg.norm(FCC_B, axis=1)[:, None]
TYPES = "SCXGHL"  # self, coplanar, cross slip, glissile, Hirth, Lomer-Cottrell


def junction_type(i, j):
    """Arsenlis-Parks class of slip systems i, j from the geometry alone."""
    if i == j:
        return "S"
    if abs(abs(FCC_N[i] @ FCC_N[j]) - 3.0) < 1e-12:
        return "C"
    bb = FCC_B[i] @ FCC_B[j]
    if abs(abs(bb) - 2.0) < 1e-12:
        return "X"
    if abs(bb) < 1e-12:
        return "H"
    # the junction's Burgers vector b_i + b_j or b_i - b_j (the <110> one):
    # glissile when it lies in one of the two planes, a Lomer-Cottrell lock
    # when it lies in neither
    br = FCC_B[i] - np.sign(bb) * FCC_B[j]
    return "G" if min(abs(FCC_N[i] @ br), abs(FCC_N[j] @ br)) < 1e-12 else "L"


def test_interaction_matrix_matches_slip_geometry():
    # distinct strengths per class, so any misplaced entry shows
    inter = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    chi = O.dd_interaction(inter)
    want = np.array([[inter[TYPES.index(junction_type(i, j))] for j in range(12)]
                     for i in range(12)])
    np.testing.assert_array_equal(chi, want)
    # the class counts of one FCC system: 1 self, 2 coplanar, 1 cross slip,
    # 2 Hirth, 4 glissile, 2 Lomer-Cottrell
    row = [junction_type(0, j) for j in range(12)]
    assert {t: row.count(t) for t in TYPES} == dict(S=1, C=2, X=1, G=4, H=2, L=2)


def test_projection_tables_closed_forms():
    cth = O.dd_angles(N_HAT, B_HAT)
    cos_s, cos_g1, cos_g2, cos_g3, sin_s, sin_g1, sin_g2, sin_g3 = cth
    r89 = np.sqrt(8.0) / 3.0  # |n_i x n_j| of two different {111} planes
    for i in range(12):
        for j in range(12):
            same_plane = abs(abs(FCC_N[i] @ FCC_N[j]) - 3.0) < 1e-12
            # n_i . (n_j x b_j) = b_j . (n_i x n_j): zero for one plane; else
            # |n_i x n_j| times |cos| of b_j and the planes' intersection line
            # (two <110> of a {111} plane: 1 along it, 1/2 otherwise)
            if same_plane:
                assert cos_s[i, j] == pytest.approx(0.0, abs=1e-15)
            elif abs(FCC_N[i] @ FCC_B[j]) < 1e-12:  # b_j on the intersection line
                assert cos_s[i, j] == pytest.approx(r89, rel=1e-14)
            else:
                assert cos_s[i, j] == pytest.approx(r89 / 2.0, rel=1e-14)
            # {111} normal vs <110>: perpendicular or cos = sqrt(2/3)
            in_plane = abs(FCC_N[i] @ FCC_B[j]) < 1e-12
            assert cos_g1[i, j] == pytest.approx(0.0 if in_plane else np.sqrt(2.0 / 3.0),
                                                 abs=1e-15)
            # two {111} normals: parallel or cos = 1/3
            assert cos_g3[i, j] == pytest.approx(1.0 if same_plane else 1.0 / 3.0, rel=1e-14)
    np.testing.assert_array_equal(cos_g2, cos_s)
    np.testing.assert_array_equal(sin_g2, sin_s)
    for c, s in ((cos_s, sin_s), (cos_g1, sin_g1), (cos_g3, sin_g3)):
        np.testing.assert_allclose(c * c + s * s, 1.0, rtol=1e-14)
        assert (s >= 0).all() and (c >= 0).all()
    # the table is not symmetric: rows are the plane, columns the system
    assert np.abs(cos_s - cos_s.T).max() > 0.4


@pytest.mark.parametrize("ec", [(241.1674e3, 89.1989e3, 151.9685e3), (168.4e3, 121.4e3, 75.4e3)])
def test_shear_modulus_is_hill_average(ec):
    c11, c12, c44 = ec
    g_voigt = (c11 - c12 + 3.0 * c44) / 5.0
    g_reuss = 5.0 * (c11 - c12) * c44 / (4.0 * c44 + 3.0 * (c11 - c12))
    assert O.dd_shear_modulus(ec) == pytest.approx((g_voigt + g_reuss) / 2.0, rel=1e-14)
    # isotropic: c11 - c12 = 2 c44 gives exactly c44
    assert O.dd_shear_modulus([3.0e5, 1.0e5, 1.0e5]) == pytest.approx(1.0e5, rel=1e-15)


def _rate_setup(seed=1):
    rng = np.random.default_rng(seed)
    tran = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    schm_xt = np.array([O.t2_to_v5((np.outer(n, b) + np.outer(b, n)) / 2.0)
                        for n, b in zip(N_HAT, B_HAT)])
    sc = O.rotate_schmid(tran, schm_xt)
    law = synth.DDLaw()
    kT = O.dd_kT(law.T_K)
    G = O.dd_shear_modulus([241.1674e3, 89.1989e3, 151.9685e3])
    args = dict(c1=law.c[0], c2=law.c[1], c3=law.c[2], G=G,
                q_slip=law.q_slip * 0.62415063631 / kT, kT=kT, drx=0.0,
                bbB=np.full(12, law.a0 / np.sqrt(2.0)), chi=O.dd_interaction(law.inter),
                cth=O.dd_angles(N_HAT, B_HAT))
    rho = dict(rho_s=rng.uniform(0.05, 0.5, 12), rho_g1=rng.normal(0, 0.1, 12),
               rho_g2=rng.normal(0, 0.1, 12), rho_g3=np.zeros(12))
    return sc, args, rho, rng


def _rate(stress6, sc, args, rho, rho_m):
    return O.strain_rate_dd(stress6, sc, rho_m=rho_m, **args, **rho)


def test_dd_rate_derivative_matches_finite_differences():
    sc, args, rho, rng = _rate_setup()
    s = np.zeros(6)
    s[:5] = rng.normal(0, 150.0, 5)
    # the first call sets the mobile density; later calls see it unchanged
    st, _, _, _, rho_m, _, _ = _rate(s, sc, args, rho, np.zeros(12))
    assert st == 0
    st, e0, d0, _, rho_m2, _, _ = _rate(s, sc, args, rho, rho_m)
    np.testing.assert_array_equal(rho_m2, rho_m)
    h = 1e-4
    fd = np.zeros((6, 6))
    for j in range(5):
        sp, sm = s.copy(), s.copy()
        sp[j] += h
        sm[j] -= h
        fd[:, j] = (_rate(sp, sc, args, rho, rho_m)[1] - _rate(sm, sc, args, rho, rho_m)[1]) / (2 * h)
    assert np.abs(d0[:5, :5]).max() > 0
    np.testing.assert_allclose(d0[:5, :5], fd[:5, :5], rtol=1e-6, atol=1e-9 * np.abs(d0).max())
    assert (d0[5] == 0).all() and (d0[:, 5] == 0).all()
    # d(edot)/d(sigma) is a sum of s s^T with non-negative weights
    assert np.linalg.eigvalsh(d0[:5, :5]).min() > -1e-12 * np.abs(d0).max()


def test_dd_rate_is_odd_and_has_a_threshold():
    sc, args, rho, rng = _rate_setup(2)
    s = np.zeros(6)
    s[:5] = rng.normal(0, 150.0, 5)
    _, _, _, _, rho_m, _, _ = _rate(s, sc, args, rho, np.zeros(12))
    _, ep, _, gp, _, _, _ = _rate(s, sc, args, rho, rho_m)
    _, em, _, gm, _, _, _ = _rate(-s, sc, args, rho, rho_m)
    np.testing.assert_array_equal(ep, -em)
    np.testing.assert_array_equal(gp, -gm)
    # below every system's passing stress nothing slips (rss ~ 1e-3 MPa)
    _, e0, d0, g0, _, _, _ = _rate(s * 1e-5, sc, args, rho, rho_m)
    assert (g0 == 0).all() and (e0 == 0).all() and (d0 == 0).all()
    # the hydrostatic B-basis component never resolves onto a slip system
    h = np.zeros(6)
    h[5] = 1e4
    _, eh, _, gh, _, _, _ = _rate(h, sc, args, rho, rho_m)
    assert (gh == 0).all()


def test_dd_mobile_density_lags_one_call():
    sc, args, rho, rng = _rate_setup(3)
    s = np.zeros(6)
    s[:5] = rng.normal(0, 60.0, 5)
    st, e_a, _, _, rm_a, rF, rP = _rate(s, sc, args, rho, np.zeros(12))
    st, e_b, _, _, rm_b, _, _ = _rate(s, sc, args, rho, rm_a)
    # the mobile density depends on the densities only, so the second call
    # leaves it; the first call's passing stress used the zero it was given
    np.testing.assert_array_equal(rm_a, rm_b)
    assert (rm_a > 0).all()
    assert np.abs(e_a - e_b).max() > 0
    # forest and parallel densities are the interaction-weighted projections
    # of the (non-negative) densities, so both are positive here
    assert (rF > 0).all() and (rP > 0).all()
    # rho_m ~ sqrt(rho_P rho_F): scaling every density by 4 scales it by 4
    rho4 = {k: 4.0 * v for k, v in rho.items()}
    _, _, _, _, rm4, _, _ = O.strain_rate_dd(s, sc, rho_m=np.zeros(12), **args, **rho4)
    np.testing.assert_allclose(rm4, 4.0 * rm_a, rtol=1e-14)


def _dd_job(tmp_path, **kw):
    cfg = synth.Config(n=8, grains=8, increments=3, dd=synth.DDLaw(), **kw)
    return O.Job(synth.write_job(str(tmp_path), cfg), O.LAW_DD)


def test_dd_job_setup_tables(tmp_path):
    J = _dd_job(tmp_path)
    t = J.dd_tables()
    assert t["bbB"][0] == pytest.approx(np.full(12, 0.362 / np.sqrt(2.0)), rel=1e-15)
    # sense vectors b x n: unit, orthogonal to both
    bp = t["bp"][0]
    np.testing.assert_allclose(np.linalg.norm(bp, axis=1), 1.0, rtol=1e-14)
    np.testing.assert_allclose(np.einsum("ij,ij->i", bp, N_HAT), 0.0, atol=1e-15)
    np.testing.assert_allclose(np.cross(B_HAT, N_HAT), bp, atol=1e-15)
    # DD_CU: the copper constants replace the namelist's (src/init.c:821-826)
    c066, _, _, _ = J.setup_info()
    assert t["shear_G"][0] == pytest.approx(O.dd_shear_modulus([241.1674e3, 89.1989e3,
                                                                151.9685e3]), rel=1e-15)
    np.testing.assert_array_equal(J.field(O.RHO_S), 0.05)
    np.testing.assert_array_equal(J.field(O.RHO_M), 0.0)


def test_dd_gradient_stencil(tmp_path):
    """One-sided differences at grain boundaries, central ones inside, and the
    z test against the -y neighbour's grain (src/constitutive.c:818)."""
    cfg = synth.Config(n=8, grains=8, increments=1, dd=synth.DDLaw())
    inp = synth.write_job(str(tmp_path), cfg)
    # two grains split along y: y < 4 grain 1, y >= 4 grain 2
    n = 8
    x, y, z = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    grain = np.where(y < 4, 1, 2).ravel()
    synth.write_ms(os.path.join(str(tmp_path), "ms.txt"), n, np.zeros((n ** 3, 3)), grain,
                   np.ones(n ** 3, dtype=int))
    J = O.Job(inp, O.LAW_DD)
    f = (z.astype(float) ** 2 + 3.0 * x)  # quadratic in z, linear in x
    gd = np.repeat(f.reshape(-1, 1), 12, axis=1)
    J.set_field(O.GAMDOT, gd)
    g = J.dd_gradient()[:, 0, :].reshape(n, n, n, 3)
    zp, zm = (z + 1) % n, (z - 1) % n
    fz_p = zp ** 2 + 3.0 * x
    fz_m = zm ** 2 + 3.0 * x
    central = (fz_p - fz_m) / 2.0
    forward = fz_p - f
    # voxels whose -y neighbour is in the other grain (y = 0 and y = 4) take
    # the forward z difference; the others the central one
    other = (y == 0) | (y == 4)
    np.testing.assert_array_equal(g[..., 2][~other], central[~other])
    np.testing.assert_array_equal(g[..., 2][other], forward[other])
    # x: one grain along x everywhere, periodic central difference
    xc = ((z ** 2 + 3.0 * ((x + 1) % n)) - (z ** 2 + 3.0 * ((x - 1) % n))) / 2.0
    np.testing.assert_array_equal(g[..., 0], xc)
    # y: f does not vary along y; at y = 3 and y = 7 the +y neighbour is in
    # the other grain (backward difference), all zero
    np.testing.assert_array_equal(g[..., 1], 0.0)


def test_dd_job_evolves_densities(tmp_path):
    J = _dd_job(tmp_path)
    J.run(3)
    rs = J.field(O.RHO_S)
    gd = J.field(O.GAMDOT)
    assert rs.max() > 0.1 and (rs > 0).all()
    # systems that never slipped keep rho_s apart from the tiny climb and
    # dipole terms; slipping ones gained density
    assert np.abs(gd).max() > 0.1
    # GND rates start once the gradient is formed (Hard_Flag, istep > 1)
    assert np.abs(J.field(O.RHO_G1)).max() > 0
    np.testing.assert_array_equal(J.field(O.RHO_G3), 0.0)
    # dipole and forest sums grow under slip; AvgDislocation is the voxel mean
    # of the system sums (no pores here)
    assert J.field(O.RHO_DIPOLE).min() >= 0 and J.field(O.RHO_FOREST).max() > 0
    avg = J.macro()["dd_avg"]
    g = np.sqrt(J.field(O.RHO_G1) ** 2 + J.field(O.RHO_G2) ** 2 + J.field(O.RHO_G3) ** 2)
    want = [J.field(f).sum(axis=1).mean() for f in (O.RHO_F, O.RHO_P, O.RHO_S, O.RHO_M)]
    want += [g.sum(axis=1).mean(), J.field(O.RHO_DIPOLE).sum(axis=1).mean(),
             J.field(O.RHO_FOREST).sum(axis=1).mean()]
    np.testing.assert_allclose(avg, want, rtol=1e-13)


def test_reference_test_run_slip_is_frozen(tmp_path):
    """The reference's test_run input sets Q_slip = 2e10 (1E-19 J): exp(-q_slip)
    is 0, so the DD law gives exactly zero slip there."""
    law = synth.DDLaw(q_slip=2.0e10, T_K=30.0)
    cfg = synth.Config(n=8, grains=8, increments=2, dd=law)
    J = O.Job(synth.write_job(str(tmp_path), cfg), O.LAW_DD)
    J.run(2)
    assert (J.field(O.GAMDOT) == 0).all()
    assert (J.field(O.EDOT) == 0).all()


def test_dd_gnd_rates_from_a_uniform_shear_rate_gradient(tmp_path):
    """One grain, gamdot of every system a linear field a.x along the grid
    (periodic wrap excluded by reading interior voxels only): the central
    differences give the gradient a exactly, and the first-order GND rates
    are (a . p)/(b L0) for type 1 (p = b x n) and -(a . b)/(b L0) for type 2
    (src/constitutive.c:1418-1460), integrated over one time step from 0."""
    law = synth.DDLaw(c=(0.48, 3.0, 3.0, 0.0, 0.0, 0.0, 0.0, 0.24), q_bulk=1e9)
    cfg = synth.Config(n=8, grains=1, increments=1, dd=law, single_orientation=(0.0, 0.0, 0.0))
    inp = synth.write_job(str(tmp_path), cfg)
    J = O.Job(inp, O.LAW_DD)
    n = 8
    x, y, z = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    a = np.array([0.3, -0.2, 0.5])
    f = a[0] * x + a[1] * y + a[2] * z
    J.set_field(O.GAMDOT, np.repeat(f.reshape(-1, 1), 12, axis=1))
    J.dd_gradient()
    J.dd_evolve()
    inner = ((x > 0) & (x < n - 1) & (y > 0) & (y < n - 1) & (z > 0) & (z < n - 1)).ravel()
    t = J.dd_tables()
    b = t["bbB"][0][0]
    dt, L0 = cfg.dt, law.L0
    g1 = J.field(O.RHO_G1)[inner]
    g2 = J.field(O.RHO_G2)[inner]
    want1 = dt * (np.cross(B_HAT, N_HAT) @ a) / b / L0
    want2 = -dt * (B_HAT @ a) / b / L0
    np.testing.assert_allclose(g1, np.broadcast_to(want1, g1.shape), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(g2, np.broadcast_to(want2, g2.shape), rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(J.field(O.RHO_G3), 0.0)
    # C_4..C_7 = 0 and no climb: the SSD density is left as it was
    np.testing.assert_array_equal(J.field(O.RHO_S), law.rho_ssd_initial)


@pytest.mark.parametrize("term", ["C4", "C5", "C6", "C7", "climb"])
def test_ssd_rate_terms_one_at_a_time(tmp_path, term):
    """DislocationEvolution's SSD rate (src/constitutive.c:1395-1407) with all
    coefficients but one at zero, on a uniform crystal with set stress, shear
    rates and densities (every system different), after one call, against
    each term written from its physics in numpy:
      C4  forest storage      + |gamdot| C4 sqrt(rho_F)
      C5  dynamic recovery    - |gamdot| C5 rho_s
      C6  dipole formation    + |gamdot| C6 d rho_m, d = sqrt(3) G b /
                                (16 pi (1 - nu) (|tau| + eps)) the dipole
                                capture distance (nm: G, tau in MPa, b in nm)
      C7  static annihilation - C7 rho_s^2 (the reference's 1.6e12 x 1e-9
                                density units)
      climb                   - 4e12 d (D0 / b) rho_s^2 exp(|p| b^3 / (k T)
                                - 1) exp(-Q_bulk / (k T)), with p b^3 in J
                                (MPa nm^3 = 1e-21 J), k = 1.38e-23 J/K in the
                                pressure factor as the reference writes it and
                                the Boltzmann constant of src/evp.h:84-85 in
                                the Arrhenius one (Q_bulk in 1e-19 J); p is
                                the reference's 'hydro', a third of the sum of
                                the first three basis components (a quirk)
    The dipole and forest diagnostics grow by their terms alone."""
    import crystal_np as X
    coeff = {"C4": (3, 800.0), "C5": (4, 16.0), "C6": (5, 1100.0), "C7": (6, 1.25e-4)}
    c = [0.48, 3.0, 3.0, 0.0, 0.0, 0.0, 0.0, 0.24]
    if term in coeff:
        c[coeff[term][0]] = coeff[term][1]
    T_K, q_bulk, D0 = 700.0, 1.2, (4e-5 if term == "climb" else 0.0)
    law = synth.DDLaw(c=tuple(c), T_K=T_K, q_bulk=q_bulk, D_0=D0)
    euler = (20.0, 35.0, 50.0)
    cfg = synth.Config(n=0, dims=(2, 2, 2), grains=1, increments=1, dd=law, single_orientation=euler)
    J = O.Job(synth.write_job(str(tmp_path), cfg), O.LAW_DD)
    N = 8
    rng = np.random.default_rng(3)
    s = rng.standard_normal((3, 3)) * 60.0
    sig = s + s.T + np.eye(3) * 45.0
    gd = rng.uniform(-2.0, 2.0, 12)
    rs, rm, rf = rng.uniform(0.5, 3.0, 12), rng.uniform(0.1, 2.0, 12), rng.uniform(1.0, 9.0, 12)
    J.set_field(O.SIG, np.broadcast_to(sig, (N, 3, 3)))
    J.set_field(O.GAMDOT, np.broadcast_to(gd, (N, 12)))
    J.set_field(O.RHO_S, np.broadcast_to(rs, (N, 12)))
    J.set_field(O.RHO_M, np.broadcast_to(rm, (N, 12)))
    J.set_field(O.RHO_F, np.broadcast_to(rf, (N, 12)))
    dip0, for0 = J.field(O.RHO_DIPOLE), J.field(O.RHO_FOREST)
    J.dd_evolve()
    d_rs = J.field(O.RHO_S)[0] - rs
    d_dip = J.field(O.RHO_DIPOLE)[0] - dip0[0]
    d_for = J.field(O.RHO_FOREST)[0] - for0[0]
    # numpy: resolved shear stress from the slip geometry and the rotation
    T = X.rotation(euler)
    P = 0.5 * (np.einsum("si,sj->sij", B_HAT @ T.T, N_HAT @ T.T) +
               np.einsum("si,sj->sij", N_HAT @ T.T, B_HAT @ T.T))
    tau = np.einsum("sij,ij->s", P, sig)
    Bv = np.einsum("iab,ab->i", O.basis(), sig)
    p = (Bv[0] + Bv[1] + Bv[2]) / 3.0
    tab = J.dd_tables()
    G, b = tab["shear_G"][0], tab["bbB"][0]
    nu = synth.Material().nu
    dcap = np.sqrt(3.0) * G * b / (16.0 * X.PI_REF * (1.0 - nu)) / (np.abs(tau) + 5e-3)
    dt = cfg.dt
    ag = np.abs(gd)
    want = {
        "C4": ag * 800.0 * np.sqrt(rf),
        "C5": -ag * 16.0 * rs,
        "C6": ag * 1100.0 * dcap * rm,
        "C7": -1.25e-4 * 1.6e3 * rs ** 2,
        "climb": -4e12 * dcap * (D0 / b) * rs ** 2 *
                 np.exp(abs(p) * 1e6 * (b * 1e-9) ** 3 / (1.38e-23 * T_K) - 1.0) *
                 np.exp(-q_bulk * 1e-19 / (1.380658e-23 * T_K)),
    }[term] * dt
    assert np.all(np.abs(want) > 1e-10 * rs)  # every system's term is resolved
    np.testing.assert_allclose(d_rs, want, rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(d_dip, ag * (1100.0 if term == "C6" else 0.0) * dcap * rm * dt,
                               rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(d_for, ag * (800.0 if term == "C4" else 0.0) * np.sqrt(rf) * dt,
                               rtol=1e-12, atol=1e-300)


def test_orowan_rate_from_its_physics():
    """StrainRate_Orowan (src/constitutive.c:507-721, FORWARD_PLUS_BACKWARD)
    evaluated in numpy from its physical statement in SI units, against the
    oracle for one voxel with every system's densities different:
      rho_F, rho_P   interaction-weighted projections of the densities;
      tau_pass = c1 G b sqrt(rho_P + rho_m)   (Taylor passing stress, with the
                 mobile density of the previous call);
      tau_cut  = Q' sqrt(rho_F) / (c2 c3 b^2)  with Q' = kT (the energy that
                 sets the cutting stress of the thermally activated jump);
      rho_m    = 2 kT sqrt(rho_P rho_F) / (c1 c2 c3 G b^3);
      gam0     = rho_m nu b c2 exp(-Q_slip / kT) / sqrt(rho_F), nu = 1e10 /s;
      gamdot   = gam0 sinh((|tau| - tau_pass) / (1000 MPa + tau_cut)) sgn tau
                 above tau_pass, 0 below.
    Densities in 1e12/m^2, G and stresses in MPa, b in nm, k the reference's
    Boltzmann constant (src/evp.h:84-85): the unit conversions are done here
    in SI, so each of the reference's folded factors (1e-3, 0.1, 160.2177,
    1e-3) is checked, not copied.  160.2177 is the electron volt in 1e-21 J
    truncated to 7 digits: the comparison is at 1e-6."""
    sc, args, rho, rng = _rate_setup(4)
    s = np.zeros(6)
    s[:5] = rng.normal(0, 40.0, 5)
    rm_prev = rng.uniform(0.01, 0.2, 12)
    st, edot, _, gd, rho_m, rF, rP = _rate(s, sc, args, rho, rm_prev.copy())
    assert st == 0
    # SI
    kB, eV = 1.380658e-23, 1.60217733e-19
    T = synth.DDLaw().T_K
    kT = kB * T                                    # J
    G = args["G"] * 1e6                            # Pa
    b = args["bbB"] * 1e-9                         # m
    c1, c2, c3 = args["c1"], args["c2"], args["c3"]
    chi, cth = args["chi"], args["cth"]
    rs, g1, g2, g3 = (rho[k] for k in ("rho_s", "rho_g1", "rho_g2", "rho_g3"))
    F = np.array([np.sum(chi[i] * (rs * cth[0, i] + np.abs(g1 * cth[1, i]) + np.abs(g2 * cth[2, i]) +
                                   np.abs(g3 * cth[3, i]))) for i in range(12)])
    P = np.array([np.sum(chi[i] * (rs * cth[4, i] + np.abs(g1 * cth[5, i]) + np.abs(g2 * cth[6, i]) +
                                   np.abs(g3 * cth[7, i]))) for i in range(12)])
    np.testing.assert_allclose(rF, F, rtol=1e-14)
    np.testing.assert_allclose(rP, P, rtol=1e-14)
    F_si, P_si, m_prev_si = F * 1e12, P * 1e12, rm_prev * 1e12   # 1/m^2
    tau_pass = c1 * G * b * np.sqrt(P_si + m_prev_si) / 1e6      # MPa
    tau_cut = kT * np.sqrt(F_si) / (c2 * c3 * b ** 2) / 1e6      # MPa
    m_si = 2.0 * kT * np.sqrt(P_si * F_si) / (c1 * c2 * c3 * G * b ** 3)
    np.testing.assert_allclose(rho_m, m_si / 1e12, rtol=1e-6)
    q = synth.DDLaw().q_slip * 1e-19 / kT
    gam0 = m_si * 1e10 * b * c2 * np.exp(-q) / np.sqrt(F_si)   # 1/s
    tau = sc @ s[:5]
    x = (np.abs(tau) - tau_pass) / (1000.0 + tau_cut)
    want = np.where(np.abs(tau) > tau_pass, gam0 * np.sinh(x) * np.sign(tau), 0.0)
    assert np.sum(np.abs(tau) > tau_pass) >= 3 and np.sum(np.abs(tau) <= tau_pass) >= 1
    np.testing.assert_allclose(gd, want, rtol=1e-6, atol=1e-12 * np.abs(want).max())
    np.testing.assert_allclose(edot[:5], sc.T @ want, rtol=1e-6, atol=1e-12 * np.abs(want).max())
