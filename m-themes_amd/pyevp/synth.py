"""Seeded synthetic EVPFFT inputs in the reference's own file formats.

Writes the three files a ``forge input.in`` run reads:

* ``input.in``  -- the namelist parsed by ``GetNameList`` (src/io.c:96-164);
* ``Cu.sx``     -- the single-crystal file parsed by ``PlasticInit`` at fixed line
  positions (src/init.c:52-231); one slip mode, so the hardening matrix comes from
  the mode that is used (the reference takes it from the *last* mode, Q9);
* ``ms.txt``    -- the microstructure read by ``InitMicroStruct``
  (src/init.c:253-266): ``phi1 Phi phi2 x y z grain phase`` per line, Bunge angles in
  degrees, 1-based x-major indices with z fastest.

Randomness comes from SplitMix64 (Steele, Lea & Flood 2014) with a recorded 64-bit
seed, so a configuration is reproduced bit-for-bit from its seed.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Optional

import numpy as np

MASK64 = (1 << 64) - 1

# The 12 {111}<110> systems, written as the reference's Cu_p.sx lists them
# (test_run/Cu_p.sx:38-49): plane normal n, then slip direction b.
FCC_SLIP = [
    ((1, 1, -1), (0, 1, 1)), ((1, 1, -1), (1, 0, 1)), ((1, 1, -1), (1, -1, 0)),
    ((1, -1, -1), (0, 1, -1)), ((1, -1, -1), (1, 0, 1)), ((1, -1, -1), (1, 1, 0)),
    ((1, -1, 1), (0, 1, 1)), ((1, -1, 1), (1, 0, -1)), ((1, -1, 1), (1, 1, 0)),
    ((1, 1, 1), (0, 1, -1)), ((1, 1, 1), (1, 0, -1)), ((1, 1, 1), (1, -1, 0)),
]


class SplitMix64:
    """SplitMix64 generator; ``uniform()`` returns 53-bit doubles in [0, 1)."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def uniforms(self, n: int) -> np.ndarray:
        return np.array([self.uniform() for _ in range(n)], dtype=np.float64)


@dataclasses.dataclass
class Material:
    """Cubic elastic constants (MPa) and power-law/Voce parameters of one phase."""

    c11: float = 168400.0
    c12: float = 121400.0
    c44: float = 75400.0
    nu: float = 0.35
    nrs: float = 10.0       # rate exponent n
    gam0: float = 1.0       # reference shear rate (1/s)
    tau0f: float = 16.0     # initial CRSS, forward (MPa)
    tau0b: float = 16.0     # initial CRSS, backward (MPa)
    tau1: float = 10.0      # Voce tau1 (MPa)
    thet0: float = 200.0    # initial hardening rate (MPa)
    thet1: float = 20.0     # final hardening rate (MPa)
    hself: float = 1.0
    hlat: float = 1.0


@dataclasses.dataclass
class DDLaw:
    """Namelist keys of the dislocation-density law (src/io.c:47-93), phase I
    values.  Defaults follow the reference's test_run/input.in, except the
    temperature and the activation energies (1E-19 J), which are chosen so
    that slip and climb are active at the stresses of these jobs."""

    a0: float = 0.362                     # lattice parameter (nm)
    L0: float = 0.003                     # voxel length of the GND rates (mm)
    gnd_scale_ratio: float = 6000.0
    T_K: float = 300.0
    q_slip: float = 0.2
    q_bulk: float = 1.4
    c: tuple = (0.48, 3.0, 3.0, 800.0, 16.0, 1100.0, 1.25e-8, 0.24)  # C_1 .. C_8
    # Selfinter, Coplanar, CrossSlip, GlissileJunction, HirthLock, LomerCottrellLock
    inter: tuple = (0.125, 0.125, 0.725, 0.125, 0.05, 0.18)
    rho_ssd_initial: float = 0.05         # 1E12/m^2
    D_0: float = 4e-5
    tau_drx: Optional[float] = 1000.0     # None: key absent (reads as 0)
    delta_tau_drx: float = 0.0


@dataclasses.dataclass
class Config:
    """One synthetic EVPFFT job."""

    n: int = 32
    grains: int = 64
    seed: int = 0x5EED0001
    increments: int = 10
    dt: float = 1e-3
    err: float = 1e-6
    iter_max: int = 2000
    pore_fraction: float = 0.0
    pore_radius: int = 3
    material: Material = dataclasses.field(default_factory=Material)
    update_flag: int = 1
    hard_flag: int = 1
    elast_bc: int = 1
    velgrad_bc_flag: tuple = (0, 0, 1, 1, 1, 1)
    velgrad_bc: tuple = (0.0, 0.0, 1.0, 0.0, 0.0, 0.0)
    stress_bc_flag: tuple = (1, 1, 0, 0, 0, 0)
    stress_bc: tuple = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    single_orientation: Optional[tuple] = None  # (phi1, Phi, phi2) for every grain
    # a second crystalline phase (own .sx file and elastic constants) on the
    # even-numbered grains; excludes pores (the reference has two phases)
    phase2: Optional[Material] = None
    # (nx, ny, nz) of a non-cubic grid; None: n^3
    dims: Optional[tuple] = None
    # dislocation-density law keys; None: the namelist holds none
    dd: Optional[DDLaw] = None

    def shape(self) -> tuple:
        return tuple(int(d) for d in self.dims) if self.dims is not None else (self.n,) * 3


# The configurations the loop is measured and checked on (sizes, grains, loading).
def config_c1() -> Config:
    """32^3 periodic voxels, 64 Voronoi grains, 10 increments."""
    return Config(n=32, grains=64, seed=0x5EED0001, increments=10)


def config_c2() -> Config:
    """128^3, 512 grains, 2 % spherical pores of radius 3, 50 increments."""
    return Config(n=128, grains=512, seed=0x5EED0002, increments=50, pore_fraction=0.02)


def config_c3() -> Config:
    """256^3, 4096 grains, 20 increments."""
    return Config(n=256, grains=4096, seed=0x5EED0003, increments=20)


def config_c4() -> Config:
    """64^3, 128 grains, 20 increments (adjoint calibration)."""
    return Config(n=64, grains=128, seed=0x5EED0004, increments=20)


def config_c5() -> Config:
    """128^3, 256 grains, 30 increments under the dislocation-density law (the
    reference's DRX configuration; its nucleation gate, e_vm > 4, stays shut
    at 3 % strain)."""
    return Config(n=128, grains=256, seed=0x5EED0005, increments=30, dd=DDLaw())


def _shape(n) -> tuple:
    return (n, n, n) if np.isscalar(n) else tuple(int(d) for d in n)


def voronoi(n, seeds: np.ndarray) -> np.ndarray:
    """Periodic Voronoi labels (0-based) of voxel points (i, j, k), x-major, z fastest.

    ``n``: the edge of a cubic grid or its (nx, ny, nz).  Nearest seed under the
    minimum-image distance; ties go to the lower seed index.
    """
    nx, ny, nz = _shape(n)
    g = seeds.shape[0]
    ids = np.empty(nx * ny * nz, dtype=np.int32)
    cy = np.arange(ny, dtype=np.float64)
    zz = np.arange(nz, dtype=np.float64)[None, :]
    for x in range(nx):
        best = np.full((ny, nz), np.inf)
        lab = np.zeros((ny, nz), dtype=np.int32)
        dx = np.abs(x - seeds[:, 0])
        dx = np.minimum(dx, nx - dx)
        for s in range(g):
            dy = np.abs(cy - seeds[s, 1])
            dy = np.minimum(dy, ny - dy)
            dz = np.abs(zz - seeds[s, 2])
            dz = np.minimum(dz, nz - dz)
            d = dx[s] * dx[s] + dy[:, None] ** 2 + dz ** 2
            m = d < best
            best = np.where(m, d, best)
            lab = np.where(m, s, lab)
        ids[x * ny * nz:(x + 1) * ny * nz] = lab.reshape(-1)
    return ids


def voronoi_fast(n, seeds: np.ndarray) -> np.ndarray:
    """Same labels as :func:`voronoi` via a periodic k-d tree (large grids)."""
    from scipy.spatial import cKDTree

    nx, ny, nz = _shape(n)
    box = np.array([nx, ny, nz], dtype=np.float64)
    tree = cKDTree(np.mod(seeds, box), boxsize=box)
    out = np.empty(nx * ny * nz, dtype=np.int32)
    yy, zz = np.meshgrid(np.arange(ny, dtype=np.float64), np.arange(nz, dtype=np.float64), indexing="ij")
    plane = np.stack([np.zeros(ny * nz), yy.ravel(), zz.ravel()], axis=1)
    for x in range(nx):
        plane[:, 0] = x
        d, i = tree.query(plane, k=2)
        # exact ties resolved towards the lower seed index, as in voronoi()
        tie = d[:, 0] == d[:, 1]
        lab = np.where(tie, np.minimum(i[:, 0], i[:, 1]), i[:, 0])
        out[x * ny * nz:(x + 1) * ny * nz] = lab
    return out


def pores(n, frac: float, radius: int, rng: SplitMix64) -> np.ndarray:
    """Boolean gas mask: non-overlapping periodic digital spheres |x-c|^2 <= r^2."""
    nx, ny, nz = _shape(n)
    mask = np.zeros((nx, ny, nz), dtype=bool)
    if frac <= 0.0:
        return mask.reshape(-1)
    r = radius
    off = np.arange(-r, r + 1)
    ox, oy, oz = np.meshgrid(off, off, off, indexing="ij")
    ball = (ox ** 2 + oy ** 2 + oz ** 2) <= r * r
    bx, by, bz = ox[ball], oy[ball], oz[ball]
    target = frac * nx * ny * nz
    tries = 0
    while mask.sum() < target and tries < 100000:
        tries += 1
        c = [int(rng.uniform() * d) for d in (nx, ny, nz)]
        px = (c[0] + bx) % nx
        py = (c[1] + by) % ny
        pz = (c[2] + bz) % nz
        # reject overlapping or touching spheres
        gx = (c[0] + np.arange(-r - 1, r + 2)) % nx
        gy = (c[1] + np.arange(-r - 1, r + 2)) % ny
        gz = (c[2] + np.arange(-r - 1, r + 2)) % nz
        if mask[np.ix_(gx, gy, gz)].any():
            continue
        mask[px, py, pz] = True
    return mask.reshape(-1)


def microstructure(cfg: Config):
    """(euler_deg[N^3,3], grain[N^3], phase[N^3]) of a configuration."""
    shape = cfg.shape()
    N = int(np.prod(shape))
    rng = SplitMix64(cfg.seed)
    seeds = rng.uniforms(3 * cfg.grains).reshape(cfg.grains, 3) * np.array(shape, dtype=np.float64)
    u = rng.uniforms(3 * cfg.grains).reshape(cfg.grains, 3)
    if cfg.single_orientation is not None:
        eul = np.tile(np.asarray(cfg.single_orientation, dtype=np.float64), (cfg.grains, 1))
    else:
        eul = np.stack([360.0 * u[:, 0], np.degrees(np.arccos(2.0 * u[:, 1] - 1.0)),
                        360.0 * u[:, 2]], axis=1)
    # round to the precision written to the file so in-memory and file inputs agree
    eul = np.round(eul, 8)
    if cfg.grains == 1:
        lab = np.zeros(N, dtype=np.int32)
    elif N * cfg.grains <= 4_000_000:
        lab = voronoi(shape, seeds)
    else:
        lab = voronoi_fast(shape, seeds)
    gas = pores(shape, cfg.pore_fraction, cfg.pore_radius, rng)
    grain = (lab + 1).astype(np.int32)
    phase = np.ones(N, dtype=np.int32)
    if cfg.phase2 is not None:
        if cfg.pore_fraction > 0.0:
            raise ValueError("a second crystalline phase and pores need three phases")
        phase[grain % 2 == 0] = 2
    grain[gas] = cfg.grains + 1
    phase[gas] = 2
    euler = eul[lab]
    euler[gas] = 0.0
    return euler, grain, phase


def write_ms(path: str, n, euler: np.ndarray, grain: np.ndarray, phase: np.ndarray):
    """Write the 8-column microstructure file (each line < 80 characters).

    Lines are ``phi1 Phi phi2 x y z grain phase`` with the angles to 8 decimals;
    when every grain carries one orientation (the synthetic case) the angle text
    is formatted once per grain.
    """
    nx, ny, nz = _shape(n)
    N = nx * ny * nz
    idx = np.arange(N)
    xs = (idx // (ny * nz) + 1).tolist()
    ys = ((idx // nz) % ny + 1).tolist()
    zs = (idx % nz + 1).tolist()
    g = grain.tolist()
    p = phase.tolist()
    uniq, first = np.unique(grain, return_index=True)
    per_grain = np.all(euler[first][np.searchsorted(uniq, grain)] == euler)
    with open(path, "w") as f:
        if per_grain:
            txt = {int(u): "%.8f %.8f %.8f" % tuple(euler[i]) for u, i in zip(uniq, first)}
            step = 1 << 18
            for a in range(0, N, step):
                b = min(N, a + step)
                f.write("".join(f"{txt[g[v]]} {xs[v]} {ys[v]} {zs[v]} {g[v]} {p[v]}\n"
                                for v in range(a, b)))
        else:
            cols = np.column_stack([euler, xs, ys, zs, grain, phase])
            np.savetxt(f, cols, fmt=["%.8f", "%.8f", "%.8f", "%d", "%d", "%d", "%d", "%d"],
                       delimiter=" ")


def write_sx(path: str, m: Material):
    """Single-mode cubic .sx file laid out on the line positions PlasticInit reads."""
    lines = [
        "# Crystal system (XtalSys):", "CUBIC", "",
        "# Crystal axex (XtalAxis):", "1.0 1.0 1.0", "",
        "# Total number of modes listed in the file (N_modes_max):", "1", "",
        "# Number of modes to be used in the calculation (N_modes):", "1", "",
        "# Labels of the modes to be used (iMode):", "1", "",
        "# Mode# 1 -- <110>(111) SLIP",
        "\tnsmx\tnrsx\tgamd0x\ttwshx\tisectwx",
        f"\t12\t\t{float(m.nrs)!r}\t\t{float(m.gam0)!r}\t\t0.0\t\t0",
        "\ttau0xf\ttau0xb\ttau1x\tthet0\tthet1",
        f"\t{float(m.tau0f)!r}\t\t{float(m.tau0b)!r}\t\t{float(m.tau1)!r}\t\t{float(m.thet0)!r}\t{float(m.thet1)!r}",
        "\thselfx\thlatex",
        f"\t{float(m.hself)!r}\t\t{float(m.hlat)!r}",
        "\tSlip (n-b)",
    ]
    for nv, bv in FCC_SLIP:
        lines.append("\t%d %d %d\t\t%d %d %d" % (*nv, *bv))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_input(path: str, cfg: Config, sx_name="Cu.sx", ms_name="ms.txt"):
    """Namelist in the layout of test_run/input.in (power-law keys, then the
    dislocation-density keys when ``cfg.dd`` is set)."""
    m = cfg.material
    m2 = cfg.phase2
    two = cfg.pore_fraction > 0.0 or m2 is not None
    lines = [
        "# synthetic EVPFFT job (seed 0x%X)" % cfg.seed,
        "N_phases\t%d" % (2 if two else 1),
        "CellDim\t%d %d %d" % cfg.shape(),
        "initial_ms\t%s" % ms_name,
        "Type_phases\t0\t%d" % (0 if m2 is not None else 1),
        "ElastConst_PhaseI\t%r\t%r\t%r\t%r" % (m.c11, m.c12, m.c44, m.nu),
        "Slip_PhaseI\t%s" % sx_name,
        ("ElastConst_PhaseII\t%r\t%r\t%r\t%r" % (m2.c11, m2.c12, m2.c44, m2.nu) if m2 is not None
         else "ElastConst_PhaseII\t0.0\t0.0\t0.0\t0.0"),
        "Slip_PhaseII\t%s" % ("P2.sx" if m2 is not None else sx_name),
        "VelGrad_BC_Flag\t" + " ".join(str(int(v)) for v in cfg.velgrad_bc_flag),
        "VelGrad_BC\t" + " ".join(repr(float(v)) for v in cfg.velgrad_bc),
        "Stress_BC_Flag\t" + " ".join(str(int(v)) for v in cfg.stress_bc_flag),
        "Stress_BC\t" + " ".join(repr(float(v)) for v in cfg.stress_bc),
        "ElastBC\t%d" % cfg.elast_bc,
        "TimeStep %r" % cfg.dt,
        "N_steps\t%d" % cfg.increments,
        "Err\t%r" % cfg.err,
        "IterMax\t%d" % cfg.iter_max,
        "PrintControl\t0 2000",
        "Update_Flag\t%d" % cfg.update_flag,
        "Hard_Flag\t%d" % cfg.hard_flag,
        "Tex_Flag\t0",
    ]
    d = cfg.dd
    if d is not None:
        lines += [
            "a0\t%r\t0.0" % d.a0, "L0\t%r" % d.L0, "GND_ScaleRatio\t%r" % d.gnd_scale_ratio,
            "Q_slip\t%r\t0.0" % d.q_slip, "Q_bulk\t%r\t0.0" % d.q_bulk,
            "rho_SSD_initial\t%r" % d.rho_ssd_initial, "T__K\t%r" % d.T_K,
        ]
        lines += ["C_%d\t%r\t0.0" % (i + 1, v) for i, v in enumerate(d.c)]
        names = ("Selfinter", "Coplanar", "CrossSlip", "GlissileJunction", "HirthLock",
                 "LomerCottrellLock")
        lines += ["%s\t%r\t0.0" % (k, v) for k, v in zip(names, d.inter)]
        lines += ["D_0\t%r" % d.D_0, "DeltaTau_DRX\t%r" % d.delta_tau_drx]
        if d.tau_drx is not None:
            lines.append("tau_DRX\t%r" % d.tau_drx)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_job(dirpath: str, cfg: Config) -> str:
    """Write input.in, Cu.sx and ms.txt of ``cfg`` into ``dirpath``; return input.in path."""
    os.makedirs(dirpath, exist_ok=True)
    euler, grain, phase = microstructure(cfg)
    write_ms(os.path.join(dirpath, "ms.txt"), cfg.shape(), euler, grain, phase)
    write_sx(os.path.join(dirpath, "Cu.sx"), cfg.material)
    if cfg.phase2 is not None:
        write_sx(os.path.join(dirpath, "P2.sx"), cfg.phase2)
    inp = os.path.join(dirpath, "input.in")
    write_input(inp, cfg)
    return inp


# ------------------------------------------------------------ phase field --
# Inputs of the phase-field recrystallisation step (PF_XRX, src/pf_xrx.c:1469):
# the grain field and stored energy it starts from.  Parameters follow the
# reference's test_run/input.in (E_gb 0.625, M_gb 7.4e-11, Scale_Fdeform 0.5,
# pf_length_scale 2.5e-6, pf_time_step 0.004); mu and bb set the stored-energy
# scale E_disl = alpha mu bb^2 (bb = 1 makes nuclei grow within a few substeps).
PF_PARAMS = dict(E_gb=0.625, M_gb=7.4e-11, mu=4.0e4, bb=1.0, alpha=0.5, pf_length_scale=2.5e-6,
                 pf_time_step=0.004)


def pf_params(init_gid: int, **kw) -> dict:
    p = dict(PF_PARAMS)
    p.update(kw)
    p["init_gid"] = int(init_gid)
    return p


def pf_inputs(shape, grains: int, nuclei: int, seed: int, radius2: int = 2):
    """Periodic Voronoi grains 1..grains (uniform seeds), stored energy
    U(200, 250), and `nuclei` spheres (squared radius radius2, periodic) of new
    grains grains+1.. with the stored energy 8 per slip system x 12 = 96 that
    the reference gives a nucleus (src/evolution.c:977-983).  Returns (grain
    field int32, stored energy, parameters with init_gid = grains)."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    shape = tuple(int(s) for s in shape)
    L = np.array(shape, float)
    seeds = rng.uniform(0, 1, (grains, 3)) * L
    pts = np.indices(shape).reshape(3, -1).T + 0.5
    _, near = cKDTree(seeds, boxsize=L).query(pts)
    gid = (near + 1).astype(np.int32).reshape(shape)
    rho = 200.0 + rng.uniform(0.0, 50.0, shape)
    idx = np.indices(shape).reshape(3, -1).T
    Li = np.array(shape)
    for k in range(nuclei):
        c = rng.integers(0, Li)
        d = np.abs(idx - c)
        d = np.minimum(d, Li - d)
        m = ((d * d).sum(1) <= radius2).reshape(shape)
        gid[m] = grains + 1 + k
        rho[m] = 96.0
    return gid, rho, pf_params(grains)
