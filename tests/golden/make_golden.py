"""Write the golden fixtures the GPU parity tests compare against.

Runs the CPU oracle (oracle/) on
  * the reference's own test_run/ inputs (input.in, Cu_p.sx, the 32^3
    pore microstructure), all N_steps = 2 increments;
  * the 32^3 synthetic job (64 Voronoi grains, seed 0x5EED0001), all 10
    increments of its configuration;
  * a 16^3 job with 2 % pores (8 grains), 20 increments: orientation updates
    and hardening compounding over a longer horizon;
and stores per-increment Green iteration counts, macroscopic stress and the
vm.out row, plus the final stress / strain-rate / displacement-gradient /
CRSS fields on every 16th voxel.  The test_run inputs are also run under the
dislocation-density law the reference's build compiles in
(oracle_test_run_dd.npz, with the densities on the same voxels).

    python tests/golden/make_golden.py        # all fixtures
    python tests/golden/make_golden.py dd     # the dislocation-density one only
    python tests/golden/make_golden.py c1     # the 32^3 job only
    python tests/golden/make_golden.py long   # the 16^3, 20-increment job only
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "m-themes_amd"))

import oracle as O  # noqa: E402
from pyevp import synth  # noqa: E402

STRIDE = 16


def record(inp, nsteps, out, law=O.LAW_POWER):
    J = O.Job(inp, law)
    iters, sig_avg, vm = [], [], []
    for _ in range(nsteps):
        iters.append(int(J.run(1)[0]))
        m = J.macro()
        sig_avg.append(m["SigAvg"].ravel())
        vm.append([m[k] for k in ("TimeTot", "e_vm", "evmp", "d_vm", "dvmp", "s_vm", "s_vm1")])
    sel = np.arange(0, J.n, STRIDE)
    extra = {}
    if law == O.LAW_DD:
        extra = dict(rho_s=J.field(O.RHO_S)[sel], rho_m=J.field(O.RHO_M)[sel],
                     rho_F=J.field(O.RHO_F)[sel], rho_P=J.field(O.RHO_P)[sel],
                     dd_avg=J.macro()["dd_avg"])
    np.savez_compressed(
        out, iters=np.array(iters), sig_avg=np.array(sig_avg), vm=np.array(vm), sel=sel,
        sig=J.field(O.SIG)[sel], edot=J.field(O.EDOT)[sel], disgrad=J.field(O.DISGRAD)[sel],
        crss=J.field(O.CRSS)[sel], tranmat=J.field(O.TRANMAT)[sel], grain=J.grain(),
        phase=J.phase(), **extra)
    print(out, iters, flush=True)


def c1():
    with tempfile.TemporaryDirectory() as d:
        inp = synth.write_job(d, synth.config_c1())
        record(inp, synth.config_c1().increments, os.path.join(HERE, "oracle_c1_32.npz"))


def long16():
    cfg = synth.Config(n=16, grains=8, seed=0x5EED0016, increments=20, pore_fraction=0.02, pore_radius=1)
    with tempfile.TemporaryDirectory() as d:
        record(synth.write_job(d, cfg), cfg.increments, os.path.join(HERE, "oracle_16_pores_20.npz"))


if __name__ == "__main__":
    which = sys.argv[1:]
    if which == ["c1"]:
        c1()
        sys.exit(0)
    if which == ["long"]:
        long16()
        sys.exit(0)
    record(os.path.join(HERE, "test_run", "input.in"), 2, os.path.join(HERE, "oracle_test_run_dd.npz"),
           O.LAW_DD)
    if which == ["dd"]:
        sys.exit(0)
    record(os.path.join(HERE, "test_run", "input.in"), 2, os.path.join(HERE, "oracle_test_run.npz"))
    c1()
    long16()
