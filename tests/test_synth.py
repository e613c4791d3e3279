"""The seeded input generator: determinism, periodic Voronoi, pores, formats."""
import numpy as np

import oracle as O
from pyevp import synth


def test_splitmix_is_deterministic_and_uniform():
    a = synth.SplitMix64(42).uniforms(2000)
    b = synth.SplitMix64(42).uniforms(2000)
    assert np.array_equal(a, b)
    assert 0.45 < a.mean() < 0.55 and a.min() >= 0 and a.max() < 1
    assert not np.array_equal(a, synth.SplitMix64(43).uniforms(2000))


def test_voronoi_kdtree_matches_brute_force():
    n = 12
    seeds = synth.SplitMix64(5).uniforms(3 * 7).reshape(7, 3) * n
    assert np.array_equal(synth.voronoi(n, seeds), synth.voronoi_fast(n, seeds))


def test_voronoi_is_periodic_nearest_seed():
    n = 10
    seeds = np.array([[0.5, 0.5, 0.5], [9.6, 9.6, 9.6]])
    lab = synth.voronoi(n, seeds).reshape(n, n, n)
    # (9,9,9) is nearer to the second seed; (0,0,0) is 0.87 from seed 0 and 0.69 from
    # the periodic image of seed 1
    assert lab[9, 9, 9] == 1 and lab[0, 0, 0] == 1 and lab[3, 3, 3] == 0


def test_pores_fraction_and_shape():
    m = synth.pores(32, 0.02, 2, synth.SplitMix64(9)).reshape(32, 32, 32)
    assert 0.02 <= m.mean() < 0.03
    # digital spheres of radius 2 hold 33 voxels; the spheres do not touch
    assert m.sum() % 33 == 0


def test_job_files_round_trip_through_the_oracle_reader(tmp_path):
    cfg = synth.Config(n=8, grains=5, seed=3, increments=1, pore_fraction=0.05, pore_radius=1)
    inp = synth.write_job(str(tmp_path), cfg)
    eul, grain, phase = synth.microstructure(cfg)
    J = O.Job(inp)
    assert np.array_equal(J.grain(), grain) and np.array_equal(J.phase(), phase)
    # every line fits the reference's 80-byte reads
    assert max(len(l) for l in open(tmp_path / "ms.txt")) < 80
    # uniform SO(3): cos(Phi) uniform in [-1, 1]
    big = synth.Config(n=8, grains=4000, seed=1)
    rng = synth.SplitMix64(big.seed)
    rng.uniforms(3 * big.grains)
    u = rng.uniforms(3 * big.grains).reshape(-1, 3)
    c = 2.0 * u[:, 1] - 1.0
    assert abs(c.mean()) < 0.05 and abs((c ** 2).mean() - 1 / 3) < 0.03


def test_two_crystal_phases_reach_the_oracle(tmp_path):
    """The second crystalline phase (even grains) is read with its own elastic
    constants and slip file, and its rate exponent changes the response."""
    m2 = synth.Material(c11=246500.0, c12=147300.0, c44=124700.0, nrs=20.0, tau0f=30.0,
                        tau0b=30.0)
    cfg = synth.Config(n=8, grains=6, seed=31, increments=1, phase2=m2)
    inp = synth.write_job(str(tmp_path), cfg)
    _, grain, phase = synth.microstructure(cfg)
    assert set(np.unique(phase)) == {1, 2}
    assert np.array_equal(phase == 2, grain % 2 == 0)
    J = O.Job(inp)
    assert np.array_equal(J.phase(), phase)
    J.run(1)
    one = synth.write_job(str(tmp_path / "one"), synth.Config(n=8, grains=6, seed=31, increments=1))
    J1 = O.Job(one)
    J1.run(1)
    assert abs(J.macro()["SigAvg"][2, 2] - J1.macro()["SigAvg"][2, 2]) > 1e-3 * abs(J1.macro()["SigAvg"][2, 2])


def test_non_cubic_grid_round_trips_through_the_oracle(tmp_path):
    """(nx, ny, nz) grids: k-d tree and brute-force Voronoi agree, the pores
    stay in the box and the oracle reads CellDim and the voxel order back."""
    shape = (12, 6, 10)
    seeds = synth.SplitMix64(4).uniforms(3 * 5).reshape(5, 3) * np.array(shape, float)
    assert np.array_equal(synth.voronoi(shape, seeds), synth.voronoi_fast(shape, seeds))
    m = synth.pores(shape, 0.03, 1, synth.SplitMix64(2))
    # radius-1 digital spheres hold 7 voxels
    assert m.shape == (720,) and 0.03 <= m.mean() < 0.05 and m.sum() % 7 == 0
    cfg = synth.Config(n=0, dims=(8, 4, 16), grains=5, seed=11, increments=1, pore_fraction=0.03,
                       pore_radius=1)
    inp = synth.write_job(str(tmp_path), cfg)
    _, grain, phase = synth.microstructure(cfg)
    J = O.Job(inp)
    assert tuple(J.dims) == (8, 4, 16)
    assert np.array_equal(J.grain(), grain) and np.array_equal(J.phase(), phase)
    J.run(1)
