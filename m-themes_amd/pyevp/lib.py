"""ctypes bindings of libevp (drop-in, include/evp.h) and libevpgpu (include/evpgpu.h).

The libraries are built in-tree (``m-themes_amd/lib``); nothing here falls back to
another implementation: a missing library or a failed call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# the in-tree build; EVP_LIB_DIR points the loader at another build of the
# same sources (for example one made with different compile options)
LIBDIR = os.environ.get("EVP_LIB_DIR") or os.path.join(PKG, "lib")
REPO = os.path.dirname(PKG)

FP64, MIXED = 0, 1
LAW_POWER, LAW_DD = 0, 1  # evp_set_law
NSYSMX = 12

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


class Phase(C.Structure):
    _fields_ = [("is_gas", C.c_int), ("nsys", C.c_int),
                ("nrs", C.c_double * NSYSMX), ("gam0", C.c_double * NSYSMX),
                ("tau", (C.c_double * 3) * NSYSMX), ("thet", (C.c_double * 2) * NSYSMX),
                ("hard", (C.c_double * NSYSMX) * NSYSMX),
                ("schm_xt", (C.c_double * 5) * NSYSMX),
                ("bn", (C.c_double * 3) * NSYSMX), ("bb", (C.c_double * 3) * NSYSMX)]


class DDPhase(C.Structure):
    """evpgpu_dd_phase: the dislocation-density law of one crystal phase."""
    _fields_ = [("c", C.c_double * 8), ("q_slip", C.c_double), ("q_bulk", C.c_double),
                ("shear_G", C.c_double), ("poisson_nu", C.c_double),
                ("bbB", C.c_double * NSYSMX), ("bp", (C.c_double * 3) * NSYSMX),
                ("chi", (C.c_double * NSYSMX) * NSYSMX),
                ("cth", ((C.c_double * NSYSMX) * NSYSMX) * 8)]


class Loading(C.Structure):
    _fields_ = [("dt", C.c_double), ("err_tol", C.c_double), ("iter_max", C.c_int),
                ("update_flag", C.c_int), ("hard_flag", C.c_int), ("elast_bc", C.c_int),
                ("velgrad_bc_flag", C.c_int * 6), ("stress_bc_flag", C.c_int * 6),
                ("velgrad_bc", C.c_double * 6), ("stress_bc", C.c_double * 6),
                ("c066", C.c_double * 36), ("s066", C.c_double * 36),
                ("c0", C.c_double * 81), ("s0", C.c_double * 81),
                ("wgt_ph1", C.c_double), ("d_vm0", C.c_double)]


class Macro(C.Structure):
    _fields_ = [("istep", C.c_int), ("time_tot", C.c_double), ("e_vm", C.c_double),
                ("evmp", C.c_double), ("d_vm", C.c_double), ("dvmp", C.c_double),
                ("s_vm", C.c_double), ("s_vm1", C.c_double), ("err_e", C.c_double),
                ("err_s", C.c_double), ("sig_avg", C.c_double * 9), ("sig_avg1", C.c_double * 9),
                ("disgrad_avg", C.c_double * 9), ("disgrad_avg_t", C.c_double * 9),
                ("eps_avg", C.c_double * 9), ("edot_avg", C.c_double * 9)]

    def as_dict(self):
        d = {k: getattr(self, k) for k in ("istep", "time_tot", "e_vm", "evmp", "d_vm", "dvmp",
                                           "s_vm", "s_vm1", "err_e", "err_s")}
        for k in ("sig_avg", "sig_avg1", "disgrad_avg", "disgrad_avg_t", "eps_avg", "edot_avg"):
            d[k] = np.array(getattr(self, k)[:]).reshape(3, 3)
        return d


class Stats(C.Structure):
    _fields_ = [("iters", C.c_int), ("fp64_iters", C.c_int), ("newton_iters", C.c_double),
                ("ls_trials", C.c_double), ("seconds", C.c_double), ("stopped", C.c_int),
                ("err_s", C.c_double), ("err_e", C.c_double)]


class AdjointStats(C.Structure):
    _fields_ = [("cg_iters", C.c_int), ("cg_max_iters", C.c_int), ("cg_rel_resid", C.c_double),
                ("seconds", C.c_double), ("replay_seconds", C.c_double)]


PARAM_NAMES = ("tau0", "tau1", "theta0", "theta1", "n")
NPARAM = 5


# every symbol each header declares (checked by the CPU tests)
EVPGPU_SYMBOLS = [
    "evpgpu_strerror", "evpgpu_create", "evpgpu_destroy", "evpgpu_set_precision",
    "evpgpu_newton_device_noalloc",
    "evpgpu_set_anchor_interval", "evpgpu_set_newton_tolerances", "evpgpu_set_newton_shortcut",
    "evpgpu_set_fused_transforms",
    "evpgpu_set_loop_batch",
    "evpgpu_set_timing_groups",
    "evpgpu_set_phases", "evpgpu_set_loading", "evpgpu_upload_microstructure",
    "evpgpu_upload_state", "evpgpu_download_state", "evpgpu_set_macro", "evpgpu_get_macro",
    "evpgpu_increment", "evpgpu_set_timing", "evpgpu_synchronize", "evpgpu_get_timing", "evpgpu_fft_r2c_host", "evpgpu_fft_c2r_host", "evpgpu_green_host",
    "evpgpu_newton_host", "evpgpu_adjoint_record", "evpgpu_adjoint_run", "evpgpu_copy_probe",
    "evpgpu_set_dd_law", "evpgpu_download_dd_state", "evpgpu_dd_averages",
    "evpgpu_pf_create", "evpgpu_pf_destroy", "evpgpu_pf_xrx", "evpgpu_pf_set_lists",
    "evpgpu_pf_steps", "evpgpu_pf_get_lists", "evpgpu_pf_get_stats",
    "evpgpu_drx_nucleate", "evpgpu_drx_weibull", "evpgpu_drx_uniform_pos",
]
EVP_SYMBOLS = [
    "GetNameList", "PrintNameList", "OpenFiles", "CloseFiles", "SetupJob", "Evolution",
    "DestroyJob", "evp_set_abort_on_error", "evp_status", "evp_set_precision", "evp_job_open",
    "evp_job_close", "evp_job_ctx", "evp_job_dims", "evp_job_nsteps", "evp_job_run",
    "evp_job_macro", "evp_job_ids", "evp_job_grain_list", "evp_job_setup_info", "evp_job_reset",
    "evp_job_vm_row", "evp_job_set_params", "evp_job_get_params", "evp_write_iout6",
    "evp_write_texture", "evp_texture_to_text", "evp_mat_to_euler", "evp_job_write_field",
    "evp_job_write_texture", "evp_set_law", "evp_job_law",
]

_gpu = None
_evp = None


class EvpError(RuntimeError):
    def __init__(self, code, what=""):
        msg = gpu().evpgpu_strerror(code).decode() if _gpu is not None else str(code)
        super().__init__(f"{what}: status {code} ({msg})")
        self.code = code


def _load(name):
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `make -C m-themes_amd` "
                           "(or __graft_entry__.build())")
    return C.CDLL(path)


def gpu():
    """libevpgpu with prototypes."""
    global _gpu
    if _gpu is None:
        L = _load("libevpgpu.so")
        VP = C.c_void_p
        L.evpgpu_strerror.restype = C.c_char_p
        L.evpgpu_strerror.argtypes = [C.c_int]
        L.evpgpu_create.argtypes = [C.POINTER(VP), C.c_int, C.c_int, C.c_int, C.c_int]
        L.evpgpu_destroy.argtypes = [VP]
        L.evpgpu_set_precision.argtypes = [VP, C.c_int]
        L.evpgpu_set_anchor_interval.argtypes = [VP, C.c_int]
        L.evpgpu_set_newton_tolerances.argtypes = [VP, C.c_double, C.c_double]
        L.evpgpu_set_newton_shortcut.argtypes = [VP, C.c_int]
        L.evpgpu_set_fused_transforms.argtypes = [VP, C.c_int]
        L.evpgpu_set_loop_batch.argtypes = [VP, C.c_int]
        L.evpgpu_set_timing_groups.argtypes = [VP, C.c_uint, C.c_int]
        L.evpgpu_set_phases.argtypes = [VP, C.POINTER(Phase), C.c_int]
        L.evpgpu_set_loading.argtypes = [VP, C.POINTER(Loading)]
        L.evpgpu_upload_microstructure.argtypes = [VP, IP, IP, IP, C.c_int, DP]
        L.evpgpu_upload_state.argtypes = [VP] + [DP] * 7
        L.evpgpu_download_state.argtypes = [VP] + [DP] * 8
        L.evpgpu_set_macro.argtypes = [VP, C.POINTER(Macro)]
        L.evpgpu_get_macro.argtypes = [VP, C.POINTER(Macro)]
        L.evpgpu_increment.argtypes = [VP, C.c_int, C.POINTER(Stats)]
        L.evpgpu_set_timing.argtypes = [VP, C.c_int]
        L.evpgpu_synchronize.argtypes = [VP]
        L.evpgpu_get_timing.argtypes = [VP, C.c_int, DP, C.POINTER(C.c_longlong)]
        L.evpgpu_fft_r2c_host.argtypes = [VP, DP, DP]
        L.evpgpu_fft_c2r_host.argtypes = [VP, DP, DP]
        L.evpgpu_green_host.argtypes = [VP, DP, DP]
        L.evpgpu_newton_host.argtypes = [VP, C.c_int, IP, IP, DP, C.c_int, DP, DP, DP, DP, DP,
                                         DP, DP, DP, IP]
        L.evpgpu_copy_probe.argtypes = [VP, C.c_longlong, C.c_int, DP]
        L.evpgpu_adjoint_record.argtypes = [VP, C.c_int, C.c_int, DP, C.POINTER(Stats)]
        L.evpgpu_adjoint_run.argtypes = [VP, DP, C.c_double, C.c_int, DP, C.POINTER(AdjointStats)]
        L.evpgpu_download_dd_state.argtypes = [VP] + [DP] * 7
        L.evpgpu_dd_averages.argtypes = [VP, DP]
        _gpu = L
    return _gpu


def evp():
    """libevp (drop-in) with prototypes."""
    global _evp
    if _evp is None:
        gpu()
        L = _load("libevp.so")
        VP = C.c_void_p
        L.evp_job_open.restype = VP
        L.evp_job_open.argtypes = [C.c_char_p, C.c_int, IP]
        L.evp_job_close.argtypes = [VP]
        L.evp_job_ctx.restype = VP
        L.evp_job_ctx.argtypes = [VP]
        L.evp_job_dims.argtypes = [VP, IP]
        L.evp_job_nsteps.argtypes = [VP]
        L.evp_job_run.argtypes = [VP, C.c_int, C.c_int, IP, C.POINTER(Stats)]
        L.evp_job_macro.argtypes = [VP, C.POINTER(Macro)]
        L.evp_job_ids.argtypes = [VP, IP, IP]
        L.evp_job_grain_list.argtypes = [VP, IP]
        L.evp_job_setup_info.argtypes = [VP, DP, DP, DP, DP]
        L.evp_job_reset.argtypes = [VP]
        L.evp_job_vm_row.argtypes = [VP, DP]
        L.evp_job_set_params.argtypes = [VP, C.c_int, DP]
        L.evp_job_get_params.argtypes = [VP, C.c_int, DP]
        L.GetNameList.argtypes = [C.POINTER(C.c_char_p)]
        L.evp_set_abort_on_error.argtypes = [C.c_int]
        L.evp_set_precision.argtypes = [C.c_int]
        L.evp_set_law.argtypes = [C.c_int]
        L.evp_job_law.argtypes = [VP]
        _evp = L
    return _evp


def _p(a, t=DP):
    return a.ctypes.data_as(t)


def _chk(st, what):
    if st:
        raise EvpError(st, what)


class GpuJob:
    """A drop-in EVPFFT job (input.in + .sx + microstructure) resident on the GPU."""

    def __init__(self, input_path: str, precision: int = FP64, law: int = LAW_POWER):
        L = evp()
        st = C.c_int(0)
        L.evp_set_law(law)
        try:
            self.h = L.evp_job_open(os.fsencode(input_path), precision, C.byref(st))
        finally:
            L.evp_set_law(LAW_POWER)
        if not self.h:
            raise EvpError(st.value, f"evp_job_open({input_path})")
        d = (C.c_int * 3)()
        L.evp_job_dims(self.h, d)
        self.dims = tuple(d)
        self.n = d[0] * d[1] * d[2]
        self.ctx = L.evp_job_ctx(self.h)

    def close(self):
        if getattr(self, "h", None):
            evp().evp_job_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def run(self, n: int, forced: int = 0):
        it = np.zeros(n, dtype=np.int32)
        stats = (Stats * n)()
        _chk(evp().evp_job_run(self.h, n, forced, _p(it, IP), stats), "evp_job_run")
        return it, [dict(iters=s.iters, fp64_iters=s.fp64_iters, newton_iters=s.newton_iters,
                         ls_trials=s.ls_trials, seconds=s.seconds, stopped=s.stopped,
                         err_s=s.err_s, err_e=s.err_e)
                    for s in stats]

    def synchronize(self):
        _chk(gpu().evpgpu_synchronize(self.ctx), "evpgpu_synchronize")

    def reset(self):
        _chk(evp().evp_job_reset(self.h), "evp_job_reset")

    def set_timing(self, on: bool, groups=None, every: int = 1):
        """HIP-event timers on or off; groups: the timer indices to record (all by
        default), in one Green iteration of every `every`."""
        _chk(gpu().evpgpu_set_timing(self.ctx, 1 if on else 0), "evpgpu_set_timing")
        if on and (groups is not None or every != 1):
            mask = 0
            for k in (range(6) if groups is None else groups):
                mask |= 1 << int(k)
            _chk(gpu().evpgpu_set_timing_groups(self.ctx, mask, int(every)), "evpgpu_set_timing_groups")

    def timing(self, which: int):
        """(summed ms, launches) of kernel group `which` (0 Newton, 1 FFT fwd, 2 Green,
        3 FFT inverse + strain update, 4 prologue/epilogue, 5 mixed-precision anchor
        refresh)."""
        ms = C.c_double(0)
        n = C.c_longlong(0)
        _chk(gpu().evpgpu_get_timing(self.ctx, which, C.byref(ms), C.byref(n)), "get_timing")
        return ms.value, n.value

    def set_precision(self, precision: int):
        _chk(gpu().evpgpu_set_precision(self.ctx, precision), "evpgpu_set_precision")

    def copy_probe(self, n: int, reps: int = 1) -> float:
        """GB/s of the library's streaming copy of n doubles (16 n bytes per launch)."""
        g = C.c_double(0)
        _chk(gpu().evpgpu_copy_probe(self.ctx, n, reps, C.byref(g)), "evpgpu_copy_probe")
        return g.value

    def set_newton_tolerances(self, tol: float = 1e-10, tolx: float = 1e-7):
        _chk(gpu().evpgpu_set_newton_tolerances(self.ctx, tol, tolx), "evpgpu_set_newton_tolerances")

    def set_newton_shortcut(self, on: bool):
        _chk(gpu().evpgpu_set_newton_shortcut(self.ctx, int(bool(on))), "evpgpu_set_newton_shortcut")

    def set_fused_transforms(self, on: bool):
        _chk(gpu().evpgpu_set_fused_transforms(self.ctx, int(bool(on))), "evpgpu_set_fused_transforms")

    def set_loop_batch(self, batch: int):
        """Green iterations per batch queued ahead of the device (0: wait for each)."""
        _chk(gpu().evpgpu_set_loop_batch(self.ctx, int(batch)), "evpgpu_set_loop_batch")

    def set_anchor_interval(self, iterations: int):
        _chk(gpu().evpgpu_set_anchor_interval(self.ctx, iterations), "evpgpu_set_anchor_interval")

    def macro(self) -> dict:
        m = Macro()
        _chk(evp().evp_job_macro(self.h, C.byref(m)), "evp_job_macro")
        return m.as_dict()

    def ids(self):
        g = np.zeros(self.n, dtype=np.int32)
        p = np.zeros(self.n, dtype=np.int32)
        _chk(evp().evp_job_ids(self.h, _p(g, IP), _p(p, IP)), "evp_job_ids")
        return g, p

    def grain_list(self):
        m = evp().evp_job_grain_list(self.h, None)
        ids = np.zeros(max(m, 1), dtype=np.int32)
        evp().evp_job_grain_list(self.h, _p(ids, IP))
        return ids[:m]

    def setup_info(self):
        c066, s066, c0, s0 = (np.zeros(k) for k in (36, 36, 81, 81))
        _chk(evp().evp_job_setup_info(self.h, _p(c066), _p(s066), _p(c0), _p(s0)), "setup_info")
        return c066.reshape(6, 6), s066.reshape(6, 6), c0.reshape(3, 3, 3, 3), s0.reshape(3, 3, 3, 3)

    # ------------------------------------------------------ calibration --
    def set_params(self, phase: int, params):
        """(tau0, tau1, theta0, theta1, n) of crystalline phase `phase` (0-based);
        resets the job to its post-setup state with the new initial CRSS."""
        q = np.ascontiguousarray(params, dtype=np.float64)
        assert q.shape == (NPARAM,)
        _chk(evp().evp_job_set_params(self.h, phase, _p(q)), "evp_job_set_params")

    def get_params(self, phase: int) -> np.ndarray:
        q = np.zeros(NPARAM)
        _chk(evp().evp_job_get_params(self.h, phase, _p(q)), "evp_job_get_params")
        return q

    def adjoint_record(self, n: int, checkpoint_every: int = 5):
        """Run n increments keeping checkpoints; returns (sig_avg[n, 6], stats list)."""
        sa = np.zeros((n, 6))
        st = (Stats * n)()
        _chk(gpu().evpgpu_adjoint_record(self.ctx, n, checkpoint_every, _p(sa), st), "adjoint_record")
        return sa, [{f: getattr(x, f) for f, _ in Stats._fields_} for x in st]

    def adjoint_run(self, sig_avg_bar, cg_tol: float = 1e-10, cg_max: int = 2000):
        """dL/d(params) [2, 5] for the seeds dL/dsig_avg [n, 6] of the recorded run."""
        b = np.ascontiguousarray(sig_avg_bar, dtype=np.float64)
        g = np.zeros((2, NPARAM))
        S = AdjointStats()
        _chk(gpu().evpgpu_adjoint_run(self.ctx, _p(b), cg_tol, cg_max, _p(g), C.byref(S)), "adjoint_run")
        return g, {f: getattr(S, f) for f, _ in AdjointStats._fields_}

    def state(self) -> dict:
        """Fields in the reference's AoS tensor layout ([N][3][3] etc.)."""
        n = self.n
        a = {k: np.zeros(m * n) for k, m in (("sig", 6), ("dg", 9), ("eps", 6), ("edot", 6),
                                              ("tran", 9), ("crss", 24), ("gac", 1), ("gd", 12))}
        _chk(gpu().evpgpu_download_state(self.ctx, _p(a["sig"]), _p(a["dg"]), _p(a["eps"]),
                                         _p(a["edot"]), _p(a["tran"]), _p(a["crss"]),
                                         _p(a["gac"]), _p(a["gd"])), "download_state")

        def full_sym(x):
            x = x.reshape(6, n).T
            t = np.empty((n, 3, 3))
            t[:, 0, 0], t[:, 1, 1], t[:, 2, 2] = x[:, 0], x[:, 1], x[:, 2]
            t[:, 1, 2] = t[:, 2, 1] = x[:, 3]
            t[:, 0, 2] = t[:, 2, 0] = x[:, 4]
            t[:, 0, 1] = t[:, 1, 0] = x[:, 5]
            return t

        return dict(Sig=full_sym(a["sig"]), Eps=full_sym(a["eps"]), Edot=full_sym(a["edot"]),
                    DisGrad=a["dg"].reshape(9, n).T.reshape(n, 3, 3),
                    TranMat=a["tran"].reshape(9, n).T.reshape(n, 3, 3),
                    crss=a["crss"].reshape(24, n).T.reshape(n, 12, 2),
                    gamacum=a["gac"].copy(), gamdot=a["gd"].reshape(12, n).T.copy())


    def dd_averages(self) -> np.ndarray:
        """AvgDislocation: rho_F, rho_P, rho_s, rho_m, |rho_g|, dipole, forest."""
        a = np.zeros(7)
        _chk(gpu().evpgpu_dd_averages(self.ctx, _p(a)), "dd_averages")
        return a

    def dd_state(self) -> dict:
        """Densities of a dislocation-density job, each [N][12] (1E12/m^2)."""
        names = ("rho_s", "rho_m", "rho_g1", "rho_g2", "rho_g3", "rho_F", "rho_P")
        a = {k: np.zeros((self.n, 12)) for k in names}
        _chk(gpu().evpgpu_download_dd_state(self.ctx, *[_p(a[k]) for k in names]), "download_dd_state")
        return a


class Context:
    """A bare evpgpu context, for the building-block entry points."""

    def __init__(self, nx, ny, nz, precision=FP64):
        h = C.c_void_p()
        _chk(gpu().evpgpu_create(C.byref(h), nx, ny, nz, precision), "evpgpu_create")
        self.h = h
        self.dims = (nx, ny, nz)

    def close(self):
        if getattr(self, "h", None):
            gpu().evpgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def rfftn(self, x: np.ndarray) -> np.ndarray:
        nx, ny, nz = self.dims
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(nx * ny * (nz // 2 + 1) * 2)
        _chk(gpu().evpgpu_fft_r2c_host(self.h, _p(x), _p(out)), "fft_r2c")
        return out.view(np.complex128).reshape(nx, ny, nz // 2 + 1)

    def irfftn_unnormalised(self, X: np.ndarray) -> np.ndarray:
        nx, ny, nz = self.dims
        X = np.ascontiguousarray(X, dtype=np.complex128)
        out = np.zeros(nx * ny * nz)
        _chk(gpu().evpgpu_fft_c2r_host(self.h, _p(X.view(np.float64)), _p(out)), "fft_c2r")
        return out.reshape(nx, ny, nz)

    def set_phases(self, phases):
        arr = (Phase * len(phases))(*phases)
        _chk(gpu().evpgpu_set_phases(self.h, arr, len(phases)), "set_phases")

    def set_loading(self, ld: Loading):
        _chk(gpu().evpgpu_set_loading(self.h, C.byref(ld)), "set_loading")

    def green(self, sig_hat6: np.ndarray) -> np.ndarray:
        """sig_hat6: [6][nx][ny][nz/2+1] complex -> e_hat [9][...] complex."""
        s = np.ascontiguousarray(sig_hat6, dtype=np.complex128)
        out = np.zeros((9,) + s.shape[1:], dtype=np.complex128)
        _chk(gpu().evpgpu_green_host(self.h, _p(s.view(np.float64)), _p(out.view(np.float64))),
             "green")
        return out

    def newton(self, phase, orient, sgr, sig6, strain6, eps6, tran9, crss24):
        """Per-voxel stress update on a batch (component-major arrays)."""
        n = len(phase)
        phase = np.ascontiguousarray(phase, dtype=np.int32)
        orient = np.ascontiguousarray(orient, dtype=np.int32)
        sgr = np.ascontiguousarray(sgr, dtype=np.float64)
        sig = np.array(sig6, dtype=np.float64, order="C")
        ed, gd = np.zeros((6, n)), np.zeros((12, n))
        err, nit = np.zeros((n, 2)), np.zeros(n, dtype=np.int32)
        args = [np.ascontiguousarray(a, dtype=np.float64) for a in (strain6, eps6, tran9, crss24)]
        st = gpu().evpgpu_newton_host(self.h, n, _p(phase, IP), _p(orient, IP), _p(sgr),
                                      sgr.size // 36, _p(sig), _p(args[0]), _p(args[1]),
                                      _p(args[2]), _p(args[3]), _p(ed), _p(gd), _p(err),
                                      _p(nit, IP))
        return st, sig, ed, gd, err, nit
