"""ctypes front end of the CPU oracle (oracle/evp_oracle.c, oracle/pf_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker of the HIP path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "lib", "liboracle.so")
_lib = None

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)

# fields of orc_job_field()
SIG, EPS, EDOT, DISGRAD, VELGRAD, TRANMAT, CRSS, GAMDOT, GAMACUM = range(9)
# dislocation-density jobs
RHO_S, RHO_M, RHO_G1, RHO_G2, RHO_G3, RHO_F, RHO_P, GADOT_GRAD, RHO_DIPOLE, RHO_FOREST = range(9, 19)
LAW_POWER, LAW_DD = 0, 1

# layout of orc_job_macro()
MACRO = dict(TimeTot=0, e_vm=1, evmp=2, d_vm=3, dvmp=4, s_vm=5, s_vm1=6, Err_e=7,
             Err_s=8, Wgt_ph1=9, green_time=55, green_iters=56, newton_iters=57,
             istep=58, status=59)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_job_setup.restype = C.c_void_p
        L.orc_job_setup.argtypes = [C.c_char_p, IP]
        L.orc_job_free.argtypes = [C.c_void_p]
        L.orc_job_run.argtypes = [C.c_void_p, C.c_int, IP, IP]
        L.orc_job_nvox.argtypes = [C.c_void_p]
        L.orc_threads.argtypes = []
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_job_dims.argtypes = [C.c_void_p, IP]
        L.orc_job_field.restype = DP
        L.orc_job_field.argtypes = [C.c_void_p, C.c_int]
        L.orc_job_grain.restype = IP
        L.orc_job_grain.argtypes = [C.c_void_p]
        L.orc_job_phase.restype = IP
        L.orc_job_phase.argtypes = [C.c_void_p]
        L.orc_job_newton_counts.restype = IP
        L.orc_job_newton_counts.argtypes = [C.c_void_p]
        L.orc_job_macro.argtypes = [C.c_void_p, DP]
        L.orc_job_setup_info.argtypes = [C.c_void_p, DP, DP, DP, DP]
        L.orc_job_cgr.argtypes = [C.c_void_p, DP]
        L.orc_job_green.argtypes = [C.c_void_p, DP, DP, DP, DP]
        L.orc_job_newton_iters.restype = C.c_double
        L.orc_job_newton_iters.argtypes = [C.c_void_p]
        L.orc_fft3d.argtypes = [DP, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_set_newton_tol.argtypes = [C.c_double, C.c_double]
        L.orc_lu_inv66.argtypes = [DP]
        L.orc_vonmises.restype = C.c_double
        L.orc_vonmises.argtypes = [DP]
        L.orc_strain_rate.argtypes = [DP, DP, DP, C.c_int, DP, DP, DP, DP, DP]
        L.orc_harden_voxel.argtypes = [C.c_int, DP, C.c_double, DP, DP, DP, DP, DP]
        L.orc_update_stress_voxel.argtypes = [DP, DP, C.c_double, DP, DP, DP, C.c_int, DP,
                                              DP, DP, DP, DP, DP, DP, DP, IP]
        L.orc_rotate_schmid.argtypes = [DP, C.c_int, DP, DP]
        L.orc_orientation.argtypes = [DP, DP]
        L.orc_euler_to_mat.argtypes = [C.c_double] * 3 + [DP]
        for nm in ("orc_v6_to_t2", "orc_t2_to_v6", "orc_c2_to_t4", "orc_t4_to_c2",
                   "orc_v5_to_t2", "orc_t2_to_v5"):
            getattr(L, nm).argtypes = [DP, DP]
        L.orc_basis.argtypes = [DP]
        L.orc_ten4_transform_inv.argtypes = [DP, DP, DP]
        L.orc_set_law.argtypes = [C.c_int]
        L.orc_job_dd_gradient.argtypes = [C.c_void_p]
        L.orc_job_dd_evolve.argtypes = [C.c_void_p]
        L.orc_job_dd_tables.argtypes = [C.c_void_p, DP, DP, DP, DP, DP]
        L.orc_dd_interaction.argtypes = [DP, DP]
        L.orc_dd_angles.argtypes = [C.c_int, DP, DP, DP]
        L.orc_dd_shear_modulus.restype = C.c_double
        L.orc_dd_shear_modulus.argtypes = [DP]
        L.orc_dd_kT.restype = C.c_double
        L.orc_dd_kT.argtypes = [C.c_double]
        L.orc_strain_rate_dd.argtypes = [DP, DP, DP, C.c_int, DP, C.c_void_p] + [DP] * 8
        PP = C.POINTER(PfParams)
        L.orc_pf_coeffs.argtypes = [PP, DP]
        L.orc_pf_xrx.argtypes = [C.c_int] * 3 + [PP, IP, DP, C.c_int, DP, C.c_int, IP, IP, IP, DP, DP]
        L.orc_pf_run_lists.argtypes = ([C.c_int] * 3 + [PP, IP, IP, DP, IP, DP, C.c_int, DP, C.c_int,
                                       IP, IP, IP, DP, DP])
        L.orc_pf_lookup.argtypes = [IP, C.c_int, IP, DP, C.c_int]
        L.orc_pf_laplacian.restype = C.c_double
        L.orc_pf_laplacian.argtypes = [DP]
        L.orc_gb_indicator.argtypes = [C.c_int] * 3 + [IP, IP]
        L.orc_drx_nucleate.argtypes = ([C.c_int] * 4 + [IP] + [DP] * 6 + [C.c_int, C.c_int, C.c_double,
                                       C.c_double, IP, IP, IP, DP, DP, C.POINTER(C.c_longlong), C.c_int, IP])
        _lib = L
    return _lib


def ptr(a: np.ndarray, t=DP):
    return a.ctypes.data_as(t)


def f64(shape):
    return np.zeros(shape, dtype=np.float64)


class OracleError(RuntimeError):
    def __init__(self, code):
        super().__init__(f"oracle status {code}")
        self.code = code


class Job:
    """A whole EVPFFT job on the CPU: setup from input.in, then increments."""

    def __init__(self, input_path: str, law: int = LAW_POWER):
        L = lib()
        st = C.c_int(0)
        L.orc_set_law(law)
        try:
            self.h = L.orc_job_setup(os.fsencode(input_path), C.byref(st))
        finally:
            L.orc_set_law(LAW_POWER)
        self.law = law
        if not self.h:
            raise OracleError(st.value)
        d = (C.c_int * 3)()
        L.orc_job_dims(self.h, d)
        self.dims = tuple(d)
        self.n = L.orc_job_nvox(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_job_free(self.h)
            self.h = None

    def run(self, n: int, forced=None) -> np.ndarray:
        it = np.zeros(n, dtype=np.int32)
        f = None if forced is None else np.ascontiguousarray(forced, dtype=np.int32)
        st = lib().orc_job_run(self.h, n, ptr(it, IP), None if f is None else ptr(f, IP))
        if st:
            raise OracleError(st)
        return it

    def field(self, which: int) -> np.ndarray:
        p = lib().orc_job_field(self.h, which)
        if not p:
            raise ValueError(f"field {which} is not held by this job")
        per = {CRSS: 24, GAMDOT: 12, GAMACUM: 1, GADOT_GRAD: 36}.get(which, 12 if which >= RHO_S else 9)
        a = np.ctypeslib.as_array(p, shape=(self.n * per,)).copy()
        if per == 36:
            return a.reshape(self.n, 12, 3)
        if per == 9:
            return a.reshape(self.n, 3, 3)
        if per == 24:
            return a.reshape(self.n, 12, 2)
        return a.reshape(self.n, per) if per > 1 else a

    def grain(self):
        return np.ctypeslib.as_array(lib().orc_job_grain(self.h), shape=(self.n,)).copy()

    def set_field(self, which: int, values: np.ndarray):
        """Overwrite a field of the job (tests of single steps)."""
        p = lib().orc_job_field(self.h, which)
        a = np.ascontiguousarray(values, dtype=np.float64).ravel()
        C.memmove(p, a.ctypes.data, a.nbytes)

    def dd_gradient(self):
        """Gradient_ShearRate of the current gamdot field (field GADOT_GRAD)."""
        if lib().orc_job_dd_gradient(self.h):
            raise ValueError("not a dislocation-density job")
        return self.field(GADOT_GRAD)

    def dd_evolve(self):
        """One DislocationEvolution call on the job's current fields."""
        if lib().orc_job_dd_evolve(self.h):
            raise ValueError("not a dislocation-density job")

    def dd_tables(self) -> dict:
        """Shear moduli, Burgers lengths, sense vectors, interaction matrices and
        projection tables of a dislocation-density job (per phase)."""
        G, b, bp = f64(2), f64((2, 12)), f64((2, 12, 3))
        chi, cth = f64((2, 12, 12)), f64((2, 8, 12, 12))
        lib().orc_job_dd_tables(self.h, ptr(G), ptr(b), ptr(bp), ptr(chi), ptr(cth))
        return dict(shear_G=G, bbB=b, bp=bp, chi=chi, cth=cth)

    def newton_counts(self):
        return np.ctypeslib.as_array(lib().orc_job_newton_counts(self.h), shape=(self.n,)).copy()

    def phase(self):
        return np.ctypeslib.as_array(lib().orc_job_phase(self.h), shape=(self.n,)).copy()

    def macro(self) -> dict:
        o = f64(72)
        lib().orc_job_macro(self.h, ptr(o))
        d = {k: float(o[v]) for k, v in MACRO.items()}
        # AvgDislocation of a dislocation-density job (src/constitutive.c:1669-1780)
        d["dd_avg"] = o[60:67].copy()
        d["SigAvg"] = o[10:19].reshape(3, 3).copy()
        d["SigAvg1"] = o[19:28].reshape(3, 3).copy()
        d["DisGradAvg"] = o[28:37].reshape(3, 3).copy()
        d["EpsAvg"] = o[37:46].reshape(3, 3).copy()
        d["EdotAvg"] = o[46:55].reshape(3, 3).copy()
        return d

    def setup_info(self):
        c066, s066, c0, s0 = f64(36), f64(36), f64(81), f64(81)
        lib().orc_job_setup_info(self.h, ptr(c066), ptr(s066), ptr(c0), ptr(s0))
        return (c066.reshape(6, 6), s066.reshape(6, 6), c0.reshape(3, 3, 3, 3),
                s0.reshape(3, 3, 3, 3))

    def cgr(self):
        a = f64(self.n * 36)
        lib().orc_job_cgr(self.h, ptr(a))
        return a.reshape(self.n, 6, 6)

    def green(self, ksr: np.ndarray, ksi: np.ndarray):
        ksr = np.ascontiguousarray(ksr, dtype=np.float64)
        ksi = np.ascontiguousarray(ksi, dtype=np.float64)
        orr, oi = np.zeros_like(ksr), np.zeros_like(ksi)
        lib().orc_job_green(self.h, ptr(ksr), ptr(ksi), ptr(orr), ptr(oi))
        return orr, oi


# ---- unit-level wrappers ----
def basis():
    B = f64((6, 3, 3))
    lib().orc_basis(ptr(B))
    return B


def _unary(name, x, out_shape):
    x = np.ascontiguousarray(x, dtype=np.float64)
    o = f64(out_shape)
    getattr(lib(), name)(ptr(x), ptr(o))
    return o


def v6_to_t2(v): return _unary("orc_v6_to_t2", v, (3, 3))
def t2_to_v6(t): return _unary("orc_t2_to_v6", t, (6,))
def v5_to_t2(v): return _unary("orc_v5_to_t2", v, (3, 3))
def t2_to_v5(t): return _unary("orc_t2_to_v5", t, (5,))
def c2_to_t4(c): return _unary("orc_c2_to_t4", c, (3, 3, 3, 3))
def t4_to_c2(t): return _unary("orc_t4_to_c2", t, (6, 6))


def euler_to_mat(t1, ph, t2):
    a = f64((3, 3))
    lib().orc_euler_to_mat(t1, ph, t2, ptr(a))
    return a


def ten4_transform_inv(a, c2):
    c1 = f64((3, 3, 3, 3))
    lib().orc_ten4_transform_inv(ptr(c1), ptr(np.ascontiguousarray(a, dtype=np.float64)),
                                 ptr(np.ascontiguousarray(c2, dtype=np.float64)))
    return c1


def lu_inv66(c):
    c = np.array(c, dtype=np.float64, order="C")
    st = lib().orc_lu_inv66(ptr(c))
    if st:
        raise OracleError(st)
    return c


def vonmises(t):
    return lib().orc_vonmises(ptr(np.ascontiguousarray(t, dtype=np.float64)))


def fft3d(x: np.ndarray, sign: int) -> np.ndarray:
    """Unnormalised 3-D DFT of a complex array with exponent sign ``sign``."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    nx, ny, nz = a.shape
    lib().orc_fft3d(a.ctypes.data_as(DP), nx, ny, nz, sign)
    return a


def strain_rate(stress6, sc, crss, nsr, gam0):
    nsys = sc.shape[0]
    edot, dedot, gd = f64(6), f64((6, 6)), f64(12)
    lib().orc_strain_rate(ptr(np.ascontiguousarray(stress6, dtype=np.float64)), ptr(edot),
                          ptr(dedot), nsys, ptr(np.ascontiguousarray(sc, dtype=np.float64)),
                          ptr(np.ascontiguousarray(crss, dtype=np.float64)),
                          ptr(np.ascontiguousarray(nsr, dtype=np.float64)),
                          ptr(np.ascontiguousarray(gam0, dtype=np.float64)), ptr(gd))
    return edot, dedot, gd[:nsys]


def harden_voxel(gamdot, dt, hard, tau, thet, crss, gamacum):
    crss = np.array(crss, dtype=np.float64, order="C")
    g = np.array([gamacum], dtype=np.float64)
    nsys = len(gamdot)
    H = f64((12, 12))
    H[:nsys, :nsys] = hard
    lib().orc_harden_voxel(nsys, ptr(np.ascontiguousarray(gamdot, dtype=np.float64)), dt,
                           ptr(H), ptr(np.ascontiguousarray(tau, dtype=np.float64)),
                           ptr(np.ascontiguousarray(thet, dtype=np.float64)), ptr(crss), ptr(g))
    return crss, float(g[0])


def update_stress_voxel(sg66, c066, dt, sig, eps, disgrad, sc, crss, nsr, gam0):
    """Returns (status, sig_out, edot, gamdot, err_s, err_e, n_iter)."""
    sig = np.array(sig, dtype=np.float64, order="C")
    edot, gd = f64((3, 3)), f64(12)
    es, ee = f64(1), f64(1)
    nit = np.zeros(1, dtype=np.int32)
    st = lib().orc_update_stress_voxel(
        ptr(np.ascontiguousarray(sg66, dtype=np.float64)),
        ptr(np.ascontiguousarray(c066, dtype=np.float64)), dt, ptr(sig),
        ptr(np.ascontiguousarray(eps, dtype=np.float64)),
        ptr(np.ascontiguousarray(disgrad, dtype=np.float64)), sc.shape[0],
        ptr(np.ascontiguousarray(sc, dtype=np.float64)),
        ptr(np.ascontiguousarray(crss, dtype=np.float64)),
        ptr(np.ascontiguousarray(nsr, dtype=np.float64)),
        ptr(np.ascontiguousarray(gam0, dtype=np.float64)), ptr(edot), ptr(gd), ptr(es),
        ptr(ee), ptr(nit, IP))
    return st, sig, edot, gd[:sc.shape[0]], float(es[0]), float(ee[0]), int(nit[0])


def rotate_schmid(tranmat, schm_xt):
    nsys = schm_xt.shape[0]
    out = f64((nsys, 5))
    lib().orc_rotate_schmid(ptr(np.ascontiguousarray(tranmat, dtype=np.float64)), nsys,
                            ptr(np.ascontiguousarray(schm_xt, dtype=np.float64)), ptr(out))
    return out


def orientation(tran, rot):
    tran = np.array(tran, dtype=np.float64, order="C")
    rot = np.array(rot, dtype=np.float64, order="C")
    lib().orc_orientation(ptr(tran), ptr(rot))
    return tran, rot


def set_newton_tol(tol: float = 1e-10, tolx: float = 1e-7):
    """Newton tolerances of every later oracle call (process-wide)."""
    lib().orc_set_newton_tol(tol, tolx)


def threads() -> int:
    return lib().orc_threads()


def set_threads(n: int):
    """Threads of the oracle's OpenMP loops (its results do not depend on it)."""
    lib().orc_set_threads(int(n))


class DDPhase(C.Structure):
    """orc_dd_phase: the rate-law constants of one crystal phase."""
    _fields_ = [("c1", C.c_double), ("c2", C.c_double), ("c3", C.c_double), ("G", C.c_double),
                ("q_slip", C.c_double), ("kT", C.c_double), ("drx", C.c_double),
                ("bbB", DP), ("chi", DP), ("cth", DP)]


def dd_interaction(inter6):
    chi = f64((12, 12))
    lib().orc_dd_interaction(ptr(np.ascontiguousarray(inter6, dtype=np.float64)), ptr(chi))
    return chi


def dd_angles(bn, bb):
    cth = f64((8, 12, 12))
    lib().orc_dd_angles(len(bn), ptr(np.ascontiguousarray(bn, dtype=np.float64)),
                        ptr(np.ascontiguousarray(bb, dtype=np.float64)), ptr(cth))
    return cth


def dd_shear_modulus(ec3):
    return lib().orc_dd_shear_modulus(ptr(np.ascontiguousarray(ec3, dtype=np.float64)))


def dd_kT(T_K: float) -> float:
    return lib().orc_dd_kT(T_K)


def strain_rate_dd(stress6, sc, c1, c2, c3, G, q_slip, kT, drx, bbB, chi, cth, rho_s, rho_g1,
                   rho_g2, rho_g3, rho_m):
    """StrainRate_Orowan for one voxel.  Returns (status, edot, dedot, gamdot,
    rho_m after the call, rho_F, rho_P)."""
    nsys = sc.shape[0]
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (bbB, chi, cth)]
    ph = DDPhase(c1, c2, c3, G, q_slip, kT, drx, ptr(keep[0]), ptr(keep[1]), ptr(keep[2]))
    rm = np.array(rho_m, dtype=np.float64, order="C")
    edot, dedot, gd, rF, rP = f64(6), f64((6, 6)), f64(12), f64(12), f64(12)
    st = lib().orc_strain_rate_dd(
        ptr(np.ascontiguousarray(stress6, dtype=np.float64)), ptr(edot), ptr(dedot), nsys,
        ptr(np.ascontiguousarray(sc, dtype=np.float64)), C.cast(C.pointer(ph), C.c_void_p),
        *[ptr(np.ascontiguousarray(a, dtype=np.float64)) for a in (rho_s, rho_g1, rho_g2, rho_g3)],
        ptr(rm), ptr(rF), ptr(rP), ptr(gd))
    return st, edot, dedot, gd[:nsys], rm, rF[:nsys], rP[:nsys]


# ---------------------------------------------------------------- phase field
PF_CAP = 24  # ORC_PF_CAP
PF_ERR_SUM, PF_ERR_CAP, PF_ERR_GRAIN = 1385, 2001, 2002
PF_PARAM_NAMES = ("E_gb", "M_gb", "mu", "bb", "alpha", "pf_length_scale", "pf_time_step")


class PfParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in PF_PARAM_NAMES] + [("init_gid", C.c_int)]


def pf_params(d: dict) -> PfParams:
    return PfParams(*[float(d[n]) for n in PF_PARAM_NAMES], int(d["init_gid"]))


def pf_coeffs(d: dict) -> np.ndarray:
    out = np.zeros(4)
    lib().orc_pf_coeffs(C.byref(pf_params(d)), ptr(out))
    return out


def pf_laplacian(a) -> float:
    a = np.ascontiguousarray(a, dtype=np.float64)
    assert a.shape == (19,)
    return lib().orc_pf_laplacian(ptr(a))


def pf_lookup(idx: list, val: list, g: int, cap: int = PF_CAP):
    """PFdata's non-const lookup on one list: (new idx, new val, position)."""
    n = C.c_int(len(idx))
    ia = np.zeros(cap, np.int32)
    va = np.zeros(cap)
    ia[:len(idx)] = idx
    va[:len(val)] = val
    k = lib().orc_pf_lookup(C.byref(n), cap, ptr(ia, IP), ptr(va), int(g))
    return list(ia[:n.value]), list(va[:n.value]), k


def _pf_out(shape, status, gid_new, n, idx, val, defeng, rho_avg):
    return dict(status=status, gid=gid_new.reshape(shape), n=n.reshape(shape),
                idx=idx.reshape(shape + (PF_CAP,)), val=val.reshape(shape + (PF_CAP,)),
                defeng=defeng.reshape(shape), rho_avg=rho_avg)


def pf_xrx(gid: np.ndarray, rho: np.ndarray, steps: int, params: dict, rho_avg: np.ndarray) -> dict:
    """PF_XRX (src/pf_xrx.c:1469-1698) on one rank: lists from the grain
    field, then `steps` substeps.  rho_avg is copied, not changed."""
    shape = tuple(gid.shape)
    g = np.ascontiguousarray(gid, dtype=np.int32)
    r = np.ascontiguousarray(rho, dtype=np.float64)
    ra = np.array(rho_avg, dtype=np.float64)
    nv = g.size
    gid_new = np.zeros(nv, np.int32)
    n = np.zeros(nv, np.int32)
    idx = np.zeros(nv * PF_CAP, np.int32)
    val = np.zeros(nv * PF_CAP)
    de = np.zeros(nv)
    st = lib().orc_pf_xrx(shape[0], shape[1], shape[2], C.byref(pf_params(params)), ptr(g, IP), ptr(r),
                          int(steps), ptr(ra), ra.size, ptr(gid_new, IP), ptr(n, IP), ptr(idx, IP),
                          ptr(val), ptr(de))
    return _pf_out(shape, st, gid_new, n, idx, val, de, ra)


def pf_run_lists(n_in, idx_in, val_in, GID, defeng, steps: int, params: dict, rho_avg) -> dict:
    """The substeps from given interior lists (n_in[v], idx_in/val_in[v, k])."""
    shape = tuple(GID.shape)
    nv = GID.size
    ni = np.ascontiguousarray(n_in, dtype=np.int32).ravel()
    ii = np.zeros((nv, PF_CAP), np.int32)
    vi = np.zeros((nv, PF_CAP))
    k = np.asarray(idx_in).reshape(nv, -1).shape[1]
    ii[:, :k] = np.asarray(idx_in).reshape(nv, k)
    vi[:, :k] = np.asarray(val_in).reshape(nv, k)
    g = np.ascontiguousarray(GID, dtype=np.int32)
    de_in = np.ascontiguousarray(defeng, dtype=np.float64)
    ra = np.array(rho_avg, dtype=np.float64)
    gid_new = np.zeros(nv, np.int32)
    n = np.zeros(nv, np.int32)
    idx = np.zeros(nv * PF_CAP, np.int32)
    val = np.zeros(nv * PF_CAP)
    de = np.zeros(nv)
    st = lib().orc_pf_run_lists(shape[0], shape[1], shape[2], C.byref(pf_params(params)), ptr(ni, IP),
                                ptr(ii, IP), ptr(vi), ptr(g, IP), ptr(de_in), int(steps), ptr(ra), ra.size,
                                ptr(gid_new, IP), ptr(n, IP), ptr(idx, IP), ptr(val), ptr(de))
    return _pf_out(shape, st, gid_new, n, idx, val, de, ra)


def gb_indicator(grain: np.ndarray) -> np.ndarray:
    g = np.ascontiguousarray(grain, dtype=np.int32)
    out = np.zeros(g.shape, np.int32)
    lib().orc_gb_indicator(*g.shape, ptr(g, IP), ptr(out, IP))
    return out


def drx_nucleate(grain, rho_s, rho_m, g1, g2, g3, kappa, n_grains, nucleus_radius, zeta_kc, k_c,
                 gid_rex, kc, max_new=100000) -> dict:
    """NucleationCheckDRX (shipped branch, first call); densities (nsys,) + grain.shape"""
    g = np.ascontiguousarray(grain, dtype=np.int32)
    arr = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho_s, rho_m, g1, g2, g3)]
    nsys = arr[0].shape[0]
    ka = np.ascontiguousarray(kappa, dtype=np.float64)
    rex = np.array(gid_rex, dtype=np.int32)
    kcv = np.array(kc, dtype=np.float64)
    out = np.zeros(g.shape, np.int32)
    nuc = np.zeros(g.shape, np.int32)
    rt = np.zeros(g.shape)
    vox = np.zeros(max(max_new, 1), np.int64)
    n_new = C.c_int(0)
    st = lib().orc_drx_nucleate(*g.shape, nsys, ptr(g, IP), *[ptr(a) for a in arr], ptr(ka), int(n_grains),
                                int(nucleus_radius), float(zeta_kc), float(k_c), ptr(out, IP), ptr(rex, IP),
                                ptr(nuc, IP), ptr(kcv), ptr(rt), vox.ctypes.data_as(C.POINTER(C.c_longlong)),
                                int(max_new), C.byref(n_new))
    return dict(status=st, grain=out, gid_rex=rex, nuclei=nuc, kc=kcv, rho_tot=rt,
                vox=vox[:min(n_new.value, max_new)], n_new=n_new.value)
