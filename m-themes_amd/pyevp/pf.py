"""Host mirror of the reference's phase-field recrystallisation step PF_XRX
(src/pf_xrx.c:1469-1698) over libevpgpu's evpgpu_pf_* entry points.

PF_XRX(L_x, L_y, L_z, NumPE, mpirank, E_gb, M_gb, mu, bb, alpha, gID, rho,
Steps, gID_new, count, Flagwrite, rho_avg) becomes PhaseField(shape).pf_xrx(
gID, rho, steps, params, rho_avg): one GPU holds the whole grid (no rank
arguments), params carries E_gb, M_gb, mu, bb, alpha and the globals the
reference reads (pf_length_scale, pf_time_step, init_gid), and the status
the reference would abort on comes back as an EvpError.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib

CAP = 16  # EVPGPU_PF_CAP
PARAM_NAMES = ("E_gb", "M_gb", "mu", "bb", "alpha", "pf_length_scale", "pf_time_step")
ERR_PF_SUM, ERR_PF_CAP, ERR_PF_GRAIN = 1385, -8, -9


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in PARAM_NAMES] + [("init_gid", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("substeps", C.c_int), ("activity_passes", C.c_int), ("trigger_passes", C.c_int),
                ("updated", C.c_longlong), ("rho_writes", C.c_longlong), ("ms", C.c_double),
                ("ms_update", C.c_double), ("ms_avg", C.c_double), ("update_launches", C.c_int)]


def params(d: dict) -> Params:
    return Params(*[float(d[n]) for n in PARAM_NAMES], int(d["init_gid"]))


_ready = False


def _lib():
    global _ready
    L = lib.gpu()
    if not _ready:
        VP, IP, DP = C.c_void_p, lib.IP, lib.DP
        PP = C.POINTER(Params)
        L.evpgpu_pf_create.argtypes = [C.POINTER(VP), C.c_int, C.c_int, C.c_int, C.c_int]
        L.evpgpu_pf_destroy.argtypes = [VP]
        L.evpgpu_pf_xrx.argtypes = [VP, PP, IP, DP, C.c_int, IP, DP]
        L.evpgpu_pf_set_lists.argtypes = [VP, IP, IP, DP, IP, DP]
        L.evpgpu_pf_steps.argtypes = [VP, PP, C.c_int, IP, DP]
        L.evpgpu_pf_get_lists.argtypes = [VP, IP, IP, DP, DP]
        L.evpgpu_pf_get_stats.argtypes = [VP, C.POINTER(Stats)]
        L.evpgpu_drx_nucleate.argtypes = ([C.c_int] * 4 + [IP] + [DP] * 6 + [C.c_int, C.c_int, C.c_double,
                                          C.c_double, IP, IP, IP, DP, DP, C.POINTER(C.c_longlong), C.c_int, IP])
        L.evpgpu_drx_weibull.argtypes = [C.c_ulong, C.c_longlong, DP, C.c_double, DP]
        L.evpgpu_drx_uniform_pos.argtypes = [C.c_ulong, C.c_longlong, DP]
        _ready = True
    return L


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class PhaseField:
    """One Lx x Ly x Lz grid of order-parameter lists resident on the GPU."""

    def __init__(self, shape, n_rho_avg: int = 60000):
        self.shape = tuple(int(s) for s in shape)
        self.nv = int(np.prod(self.shape))
        self.n_rho = int(n_rho_avg)
        h = C.c_void_p()
        lib._chk(_lib().evpgpu_pf_create(C.byref(h), *self.shape, self.n_rho), "evpgpu_pf_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            _lib().evpgpu_pf_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def pf_xrx(self, gID, rho, steps: int, prm: dict, rho_avg):
        """PF_XRX: returns (gID_new, rho_avg after the substeps)."""
        g, r = _i32(gID), _f64(rho)
        assert g.shape == self.shape and r.shape == self.shape
        ra = np.array(rho_avg, dtype=np.float64)
        assert ra.size == self.n_rho
        out = np.zeros(self.shape, np.int32)
        lib._chk(_lib().evpgpu_pf_xrx(self.h, C.byref(params(prm)), lib._p(g, lib.IP), lib._p(r), int(steps),
                                      lib._p(out, lib.IP), lib._p(ra)), "evpgpu_pf_xrx")
        return out, ra

    def set_lists(self, n, idx, val, GID, defeng):
        ni = _i32(n).reshape(self.nv)
        ii = np.zeros((self.nv, CAP), np.int32)
        vi = np.zeros((self.nv, CAP))
        k = min(CAP, np.asarray(idx).reshape(self.nv, -1).shape[1])
        ii[:, :k] = np.asarray(idx).reshape(self.nv, -1)[:, :k]
        vi[:, :k] = np.asarray(val).reshape(self.nv, -1)[:, :k]
        g, d = _i32(GID), _f64(defeng)
        lib._chk(_lib().evpgpu_pf_set_lists(self.h, lib._p(ni, lib.IP), lib._p(ii, lib.IP), lib._p(vi),
                                            lib._p(g, lib.IP), lib._p(d)), "evpgpu_pf_set_lists")

    def steps(self, prm: dict, steps: int, rho_avg):
        ra = np.array(rho_avg, dtype=np.float64)
        out = np.zeros(self.shape, np.int32)
        lib._chk(_lib().evpgpu_pf_steps(self.h, C.byref(params(prm)), int(steps), lib._p(out, lib.IP),
                                        lib._p(ra)), "evpgpu_pf_steps")
        return out, ra

    def lists(self) -> dict:
        n = np.zeros(self.nv, np.int32)
        idx = np.zeros(self.nv * CAP, np.int32)
        val = np.zeros(self.nv * CAP)
        de = np.zeros(self.nv)
        lib._chk(_lib().evpgpu_pf_get_lists(self.h, lib._p(n, lib.IP), lib._p(idx, lib.IP), lib._p(val),
                                            lib._p(de)), "evpgpu_pf_get_lists")
        return dict(n=n.reshape(self.shape), idx=idx.reshape(self.shape + (CAP,)),
                    val=val.reshape(self.shape + (CAP,)), defeng=de.reshape(self.shape))

    def stats(self) -> dict:
        s = Stats()
        lib._chk(_lib().evpgpu_pf_get_stats(self.h, C.byref(s)), "evpgpu_pf_get_stats")
        return {f: getattr(s, f) for f, _ in Stats._fields_}


def nucleate(grain, rho_s, rho_m, g1, g2, g3, kappa, n_grains: int, nucleus_radius: int, zeta_kc: float,
             k_c: float, gid_rex, kc, max_new: int = 100000) -> dict:
    """NucleationCheckDRX (src/integrated.c:161-494), the shipped branch's first
    call; densities shaped (nsys,) + grain.shape"""
    g = _i32(grain)
    arr = [_f64(a) for a in (rho_s, rho_m, g1, g2, g3)]
    nsys = arr[0].shape[0]
    ka = _f64(kappa)
    rex = np.array(gid_rex, dtype=np.int32)
    kcv = np.array(kc, dtype=np.float64)
    out = np.zeros(g.shape, np.int32)
    nuc = np.zeros(g.shape, np.int32)
    rt = np.zeros(g.shape)
    vox = np.zeros(max(max_new, 1), np.int64)
    n_new = C.c_int(0)
    lib._chk(_lib().evpgpu_drx_nucleate(*g.shape, nsys, lib._p(g, lib.IP), *[lib._p(a) for a in arr],
                                        lib._p(ka), int(n_grains), int(nucleus_radius), float(zeta_kc),
                                        float(k_c), lib._p(out, lib.IP), lib._p(rex, lib.IP),
                                        lib._p(nuc, lib.IP), lib._p(kcv), lib._p(rt),
                                        vox.ctypes.data_as(C.POINTER(C.c_longlong)), int(max_new),
                                        C.byref(n_new)), "evpgpu_drx_nucleate")
    return dict(grain=out, gid_rex=rex, nuclei=nuc, kc=kcv, rho_tot=rt, vox=vox[:n_new.value],
                n_new=n_new.value)


def weibull(seed: int, a, b: float) -> np.ndarray:
    """kappa_drx: gsl_ran_weibull(r, a[v], b) in voxel order from MT19937(seed)"""
    av = _f64(a).ravel()
    out = np.zeros(av.size)
    lib._chk(_lib().evpgpu_drx_weibull(int(seed), av.size, lib._p(av), float(b), lib._p(out)),
             "evpgpu_drx_weibull")
    return out.reshape(np.shape(a))


def uniform_pos(seed: int, n: int) -> np.ndarray:
    out = np.zeros(int(n))
    lib._chk(_lib().evpgpu_drx_uniform_pos(int(seed), int(n), lib._p(out)), "evpgpu_drx_uniform_pos")
    return out
