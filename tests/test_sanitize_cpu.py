"""The oracle (both laws) and the drop-in's host parsers under
AddressSanitizer and UndefinedBehaviorSanitizer, on the CPU: a small job with
pores, two increments each, and the namelist / .sx / DD-key parsers on the
same input."""
import os
import shutil
import subprocess

import pytest

from pyevp import synth

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_and_host_parsers_under_asan_ubsan(tmp_path):
    exe = tmp_path / "san"
    flags = ["-O1", "-g", "-std=gnu11", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-ffp-contract=off",
             "-Wno-unused-result"]
    host = ["-D_POSIX_C_SOURCE=200809L"]  # as the drop-in's Makefile builds its host code
    objs = []
    for src, extra in ((os.path.join(HERE, "san_host_main.c"), host),
                       (os.path.join(ROOT, "oracle", "evp_oracle.c"), []),
                       (os.path.join(ROOT, "m-themes_amd", "csrc", "host", "evp_io.c"), host),
                       (os.path.join(ROOT, "m-themes_amd", "csrc", "host", "evp_dd.c"), host)):
        o = tmp_path / (os.path.basename(src) + ".o")
        r = subprocess.run(["gcc", *flags, *extra, "-c", "-o", str(o), src], capture_output=True,
                           text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        objs.append(str(o))
    r = subprocess.run(["gcc", "-fsanitize=address,undefined", "-o", str(exe), *objs, "-lm"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    job = tmp_path / "job"
    inp = synth.write_job(str(job), synth.Config(n=8, grains=4, increments=2, pore_fraction=0.05,
                                                 pore_radius=1, dd=synth.DDLaw()))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe), inp, str(job)], capture_output=True, text=True, env=env,
                       timeout=600)
    assert r.returncode == 0, (r.returncode, r.stderr[-3000:])
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
    assert r.stdout.startswith("ok 12")
