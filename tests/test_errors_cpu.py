"""Failure handling of the drop-in, on the CPU: every input error the
reference turns into PError/abort becomes a status code from evp_job_open,
before any device call (src/init.c:131-133,195-197,263-266,864-890)."""
import ctypes as C
import os
import shutil

import pytest

from pyevp import lib, synth


def _open(inp):
    st = C.c_int(0)
    h = lib.evp().evp_job_open(os.fsencode(str(inp)), 0, C.byref(st))
    assert not h
    return st.value


@pytest.fixture
def job(tmp_path):
    inp = synth.write_job(str(tmp_path), synth.Config(n=8, grains=3, seed=2, increments=1))
    return tmp_path, inp


def _edit(path, old, new):
    s = open(path).read()
    assert old in s
    open(path, "w").write(s.replace(old, new))


def test_missing_input_file(tmp_path):
    assert _open(tmp_path / "nope.in") == 1


def test_two_deviatoric_strain_components_116(job):
    d, inp = job
    _edit(inp, "VelGrad_BC_Flag\t0 0 1 1 1 1", "VelGrad_BC_Flag\t1 1 0 1 1 1")
    _edit(inp, "Stress_BC_Flag\t1 1 0 0 0 0", "Stress_BC_Flag\t0 0 1 0 0 0")
    assert _open(inp) == 116


def test_inconsistent_stress_and_strain_flags_114(job):
    d, inp = job
    _edit(inp, "Stress_BC_Flag\t1 1 0 0 0 0", "Stress_BC_Flag\t1 1 1 0 0 0")
    assert _open(inp) == 114


def test_shuffled_microstructure_1139(job):
    d, inp = job
    lines = open(d / "ms.txt").read().splitlines()
    lines[5], lines[6] = lines[6], lines[5]
    open(d / "ms.txt", "w").write("\n".join(lines) + "\n")
    assert _open(inp) == 1139


def test_hardening_rates_out_of_order_230(job):
    d, inp = job
    _edit(d / "Cu.sx", "200.0\t20.0", "20.0\t200.0")
    assert _open(inp) == 230


def test_non_orthogonal_slip_system_7(job):
    d, inp = job
    _edit(d / "Cu.sx", "1 1 -1\t\t0 1 1", "1 1 -1\t\t1 1 1")
    assert _open(inp) == 7


def test_dropin_abort_exit_code(job, tmp_path):
    """Without evp_set_abort_on_error(0) the drop-in exits with the reference's code."""
    import subprocess
    d, inp = job
    _edit(inp, "Stress_BC_Flag\t1 1 0 0 0 0", "Stress_BC_Flag\t1 1 1 0 0 0")
    r = subprocess.run([os.path.join(lib.LIBDIR, "forge"), "input.in"], cwd=d,
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 114


def test_oracle_stops_on_non_finite_stress(tmp_path):
    """A CRSS so small that |tau/crss|^n overflows: the increment stops with
    an error status -- the singular-LU exit 35 when the overflow reaches the
    Newton matrix, the NaN status 9001 (the reference's isnan asserts) when
    only the error sums turn non-finite -- instead of reporting a converged
    increment of NaN fields.  (The HIP path returns 35 from its Cholesky
    pivot test, and checks the block sums after every Newton launch:
    EVPGPU_ERR_NONFINITE.)"""
    import oracle as O
    from pyevp import synth
    mat = synth.Material(tau0f=1e-35, tau0b=1e-35)
    inp = synth.write_job(str(tmp_path), synth.Config(n=8, grains=2, seed=3, increments=1, material=mat))
    J = O.Job(inp)
    with pytest.raises(O.OracleError) as e:
        J.run(1)
    assert e.value.code in (35, 9001)
