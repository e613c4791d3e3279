"""Seeded phase-field inputs for the PF_XRX tests: periodic Voronoi grains
(uniform seeds), small spherical nuclei with grain IDs above init_gid and a
lower stored energy, the way NucleationCheckDRX marks them
(src/integrated.c:393-405), and random order-parameter lists for starts away
from PF_XRX's one-entry initial state."""
from __future__ import annotations

import numpy as np

import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "m-themes_amd"))

from pyevp import synth  # noqa: E402


def params(init_gid, **kw):
    return synth.pf_params(init_gid, **kw)


def case(shape, grains, nuclei, seed, radius2=2):
    """grain field, stored energy and parameters (init_gid = grains)"""
    return synth.pf_inputs(shape, grains, nuclei, seed, radius2)


def random_lists(gid, seed, cap=16, extra=3):
    """lists that keep each voxel's grain near 1 and add up to `extra`
    neighbouring grains with values below and above the 5e-3 threshold, in
    random order (the back entry below it in about half the voxels)"""
    rng = np.random.default_rng(seed)
    shape = gid.shape
    nv = gid.size
    n = np.zeros(nv, np.int32)
    idx = np.zeros((nv, cap), np.int32)
    val = np.zeros((nv, cap))
    flat = gid.ravel()
    ng = int(gid.max())
    for v in range(nv):
        own = int(flat[v])
        others = [g for g in rng.permutation(np.arange(1, ng + 1)) if g != own][:rng.integers(0, extra + 1)]
        ent = [(own, float(rng.uniform(0.6, 1.0)))]
        for g in others:
            small = rng.random() < 0.5
            ent.append((int(g), float(rng.uniform(1e-4, 4.9e-3) if small else rng.uniform(5.1e-3, 0.3))))
        order = rng.permutation(len(ent))
        ent = [ent[i] for i in order]
        n[v] = len(ent)
        for k, (g, x) in enumerate(ent):
            idx[v, k] = g
            val[v, k] = x
    return n.reshape(shape), idx.reshape(shape + (cap,)), val.reshape(shape + (cap,))
