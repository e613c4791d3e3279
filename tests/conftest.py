"""Shared test setup: import paths, the gpu marker, synthetic job folders."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "m-themes_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
PKG = os.path.join(ROOT, "m-themes_amd")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def job_factory(tmp_path_factory):
    """Write a synthetic job (reference file formats) once per configuration."""
    from pyevp import synth

    cache = {}

    def make(name, **kw):
        if name not in cache:
            d = tmp_path_factory.mktemp(name)
            cache[name] = synth.write_job(str(d), synth.Config(**kw))
        return cache[name]

    return make
