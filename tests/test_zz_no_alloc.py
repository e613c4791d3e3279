"""evpgpu_newton_device_noalloc (the per-voxel stress update on caller-owned
device buffers, no device allocation) against evpgpu_newton_host (the same
update on host buffers, which allocates its device copies) through the C
ABI: bit-for-bit equal stress, strain rate, slip rates, error norms and
Newton counts, on two calls in a row (the second on a smaller batch, with
the scratch array still holding the first call's partial sums); the size
limit and an out-of-range orientation index are refused with their status
codes.  The test's device buffers come from hipMalloc of the HIP runtime the
library itself loads (through ctypes).""" 
import ctypes as C

import numpy as np
import pytest

import oracle as O
from pyevp import lib, synth

pytestmark = pytest.mark.gpu

VP = C.c_void_p


def _noalloc():
    L = lib.gpu()
    f = L.evpgpu_newton_device_noalloc
    f.argtypes = [VP, C.c_int, VP, VP, VP, C.c_int] + [VP] * 9
    f.restype = C.c_int
    return f


def _batch(tmp_path):
    cfg = synth.Config(n=8, grains=6, seed=21, increments=3, pore_fraction=0.05, pore_radius=1)
    inp = synth.write_job(str(tmp_path), cfg)
    J = O.Job(inp)
    J.run(2)
    rng = np.random.default_rng(9)
    sel = np.where(J.phase() == 1)[0]
    m = len(sel)
    sig, dg, eps, tran, crss = (J.field(f)[sel] for f in (O.SIG, O.DISGRAD, O.EPS, O.TRANMAT, O.CRSS))
    strain = 0.5 * (dg + dg.transpose(0, 2, 1)) + 2e-4 * rng.standard_normal((m, 3, 3))
    strain = 0.5 * (strain + strain.transpose(0, 2, 1))
    sgr = np.array([O.lu_inv66(c) for c in J.cgr()[sel]])

    def sym6(t):
        return np.stack([t[:, 0, 0], t[:, 1, 1], t[:, 2, 2], t[:, 1, 2], t[:, 0, 2], t[:, 0, 1]])

    return inp, dict(phase=np.ones(m, np.int32), orient=np.arange(m, dtype=np.int32),
                     sgr=sgr.reshape(m, 36), sig=sym6(sig), strain=sym6(strain), eps=sym6(eps),
                     tran=tran.reshape(m, 9).T.copy(), crss=crss.reshape(m, 24).T.copy())


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libevpgpu.so uses."""

    def __init__(self):
        lib.gpu()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(VP), C.c_size_t]
        self.L.hipFree.argtypes = [VP]
        self.L.hipMemcpy.argtypes = [VP, VP, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = VP()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self):
        assert self.L.hipDeviceSynchronize() == 0
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _cols(a, k):
    """The first k voxels of a component-major [c][m] array, contiguous."""
    return np.ascontiguousarray(a[:, :k]) if a.ndim == 2 else np.ascontiguousarray(a[:k])


def _run_noalloc(f, ctx, b, k, hip):
    d = {key: hip.put(_cols(b[key], k)) for key in ("phase", "orient", "sig", "strain", "eps", "tran", "crss")}
    d["sgr"] = hip.put(b["sgr"][:k])
    outs = dict(ed=np.zeros((6, k)), gd=np.zeros((12, k)), err=np.zeros((k, 2)), nit=np.zeros(k, np.int32))
    o = {key: hip.put(v) for key, v in outs.items()}
    st = f(ctx, k, d["phase"], d["orient"], d["sgr"], k, d["sig"], d["strain"], d["eps"], d["tran"], d["crss"],
           o["ed"], o["gd"], o["err"], o["nit"])
    res = (st, hip.get(d["sig"], _cols(b["sig"], k)), hip.get(o["ed"], outs["ed"]), hip.get(o["gd"], outs["gd"]),
           hip.get(o["err"], outs["err"]), hip.get(o["nit"], outs["nit"]))
    hip.free()
    return res


def test_noalloc_newton_matches_host_entry_point(tmp_path):
    f = _noalloc()
    hip = _Hip()
    inp, b = _batch(tmp_path)
    gj = lib.GpuJob(inp)
    ctx = lib.Context.__new__(lib.Context)
    ctx.h = C.c_void_p(gj.ctx)
    m = b["phase"].size
    for k in (m, m // 3):
        print(f"noalloc: batch of {k} voxels", flush=True)
        ref = lib.Context.newton(ctx, b["phase"][:k], b["orient"][:k], b["sgr"][:k], _cols(b["sig"], k),
                                 _cols(b["strain"], k), _cols(b["eps"], k), _cols(b["tran"], k),
                                 _cols(b["crss"], k))
        assert ref[0] == 0
        got = _run_noalloc(f, gj.ctx, b, k, hip)
        assert got[0] == 0
        for name, x, y in zip(("sig", "edot", "gamdot", "err", "niter"), ref[1:], got[1:]):
            assert np.array_equal(x, y), name
        assert np.max(ref[5]) >= 2  # the solves iterate
    ctx.h = None
    gj.close()


def test_noalloc_newton_refuses_bad_input(tmp_path):
    f = _noalloc()
    hip = _Hip()
    inp, b = _batch(tmp_path)
    gj = lib.GpuJob(inp)
    k = 64
    b2 = dict(b)
    b2["orient"] = b["orient"].copy()
    b2["orient"][5] = k  # one past the table of k orientations
    st = _run_noalloc(f, gj.ctx, b2, k, hip)[0]
    assert st == -1, st  # EVPGPU_ERR_ARG, before the update kernel runs
    # larger than the scratch array holds: refused before any device work
    x = hip.put(np.zeros(64))
    i = hip.put(np.zeros(64, np.int32))
    st = f(gj.ctx, (1 << 21) + 1, i, i, x, 1, *([x] * 9))
    hip.free()
    assert st == -3, st  # EVPGPU_ERR_SIZE
    # the context still works afterwards
    got = _run_noalloc(f, gj.ctx, b, k, hip)
    assert got[0] == 0
    gj.close()
