"""Grids whose lengths are not powers of two (the reference's FFTW takes any
length): the direct-DFT transform kernels against numpy, and the whole loop
against the oracle."""
import numpy as np
import pytest

import oracle as O
from pyevp import lib

pytestmark = pytest.mark.gpu

# lengths of the form 2^m 3^a 5^b divisible by 4 (12, 20, 24, 36, 60, 96, 100,
# 240, ...) take the mixed-radix Stockham kernels, the others (5, 6, 7, 9, 10,
# 15, 3) the two-level DFT ones
SIZES = [(12, 10, 9), (5, 6, 7), (24, 20, 15), (3, 8, 100), (16, 16, 12), (9, 32, 16),
         (96, 40, 36), (60, 100, 12), (240, 20, 24), (72, 144, 48)]


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("dims", SIZES)
def test_any_length_r2c_c2r_fp64_vs_numpy(dims):
    rng = np.random.default_rng(5)
    x = rng.standard_normal(dims)
    ctx = lib.Context(*dims)
    assert rel(ctx.rfftn(x), np.fft.rfftn(x)) < 1e-12
    nh = dims[2] // 2 + 1
    # arbitrary half spectrum, imaginary parts on the kz = 0 (and Nyquist) planes included
    X = rng.standard_normal(dims[:2] + (nh,)) + 1j * rng.standard_normal(dims[:2] + (nh,))
    ref = np.fft.irfftn(X, s=dims, axes=(0, 1, 2)) * np.prod(dims)
    assert rel(ctx.irfftn_unnormalised(X), ref) < 1e-12
    ctx.close()


@pytest.mark.parametrize("dims", [(12, 10, 9), (24, 20, 15), (96, 40, 36)])
def test_any_length_fp32_vs_numpy(dims):
    x = np.random.default_rng(6).standard_normal(dims)
    ctx = lib.Context(*dims, precision=lib.MIXED)
    assert rel(ctx.rfftn(x), np.fft.rfftn(x)) < 2e-6
    X = np.fft.rfftn(x)
    assert rel(ctx.irfftn_unnormalised(X), x * np.prod(dims)) < 2e-6
    ctx.close()


@pytest.mark.parametrize("dims", [(12, 10, 9), (10, 12, 12), (24, 20, 12)])
def test_loop_parity_any_size(job_factory, dims):
    """The whole loop on odd and even non-power-of-two grids, pores included,
    against the oracle (natural convergence, identical iteration counts)."""
    inp = job_factory(f"anysize{dims}", n=0, dims=dims, grains=6, seed=47, increments=3,
                      pore_fraction=0.02, pore_radius=1)
    J = O.Job(inp)
    g = lib.GpuJob(inp)
    it_o = J.run(3)
    it_g, _ = g.run(3)
    assert list(it_g) == list(it_o), (it_g, it_o)
    s = g.state()
    for k, f in (("Sig", O.SIG), ("Edot", O.EDOT), ("DisGrad", O.DISGRAD), ("crss", O.CRSS)):
        assert rel(s[k], J.field(f)) < 1e-10, k
    g.close()


def test_mixed_precision_any_size(job_factory):
    inp = job_factory("anysizemix", n=0, dims=(20, 18, 15), grains=8, seed=53, increments=2)
    J = O.Job(inp)
    it = J.run(2)
    g = lib.GpuJob(inp, lib.MIXED)
    for k in it:
        g.run(1, forced=int(k))
    assert rel(g.state()["Sig"], J.field(O.SIG)) < 1e-6
    g.close()


def test_fused_mixed_radix_vs_oracle(job_factory):
    """48 x 24 x 24: the fused plane / line kernels on mixed-radix lengths
    (24 = 3 x 8, 48 = 3 x 16), two increments forced to 40 Green iterations
    each in the oracle and on the GPU, FP64 to the loop bar and the mixed
    loop to 1e-6."""
    inp = job_factory("fusedmixed", n=0, dims=(48, 24, 24), grains=6, seed=47, increments=2,
                      pore_fraction=0.02, pore_radius=1)
    J = O.Job(inp)
    J.run(2, forced=[40, 40])
    for prec, tol in ((lib.FP64, 1e-10), (lib.MIXED, 1e-6)):
        g = lib.GpuJob(inp, prec)
        g.run(1, forced=40)
        g.run(1, forced=40)
        s = g.state()
        assert rel(s["Sig"], J.field(O.SIG)) < tol, (prec, rel(s["Sig"], J.field(O.SIG)))
        assert rel(s["DisGrad"], J.field(O.DISGRAD)) < tol
        g.close()
