"""Pure-Python model of the data-parallel phase-field substep of
m-themes_amd/csrc/pf.hip, kernel by kernel, used by the CPU tests to show
that the four facts the kernels rest on reproduce the reference's sequential
Update_grid sweep (src/pf_xrx.c:895-1311, restated in oracle/pf_oracle.c)
bit for bit.  Test infrastructure; small grids only.

Lists are Python lists of [grain, value] pairs over the padded grid (three
halo planes each side), keyed by padded (x, y, z).
"""
from __future__ import annotations

import bisect

EPS = 5.0e-3
CAP = 16

# the 9 face/edge neighbours before a voxel in the x-major sweep, closest first
EARLY = [(0, 0, -1), (0, -1, 1), (0, -1, 0), (0, -1, -1), (-1, 1, 0),
         (-1, 0, 1), (-1, 0, 0), (-1, 0, -1), (-1, -1, 0)]
# the laplacian call's neighbour cells in argument order
LAPN = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
        (-1, 0, 1), (-1, 0, -1), (-1, -1, 0), (-1, 1, 0), (0, -1, 1), (0, 1, 1),
        (0, -1, -1), (0, 1, -1), (1, 0, 1), (1, 0, -1), (1, -1, 0), (1, 1, 0)]


def laplacian(a):
    return (0.5 * (a[0] + a[1] + a[2] + a[3] + a[4] + a[5]) +
            0.125 * (a[10] + a[12] + a[11] + a[13] + a[7] + a[6] + a[8] + a[9] + a[14] + a[17] +
                     a[15] + a[16]) - 4.5 * a[18])


def coeffs(P):
    dx = P["pf_length_scale"]
    L_gb = 3.0 * dx
    dt = P["pf_time_step"]
    return (dt * 4.0 / 3.0 * P["M_gb"] / L_gb, 3.0 / 4.0 * P["E_gb"] * L_gb / dx / dx,
            6 * P["E_gb"] / L_gb, P["alpha"] * P["mu"] * P["bb"] * P["bb"])


def lookupv(L, g):
    """fact 2: the value a lookup returns from a list, from its start state"""
    if not L:
        return 0.0
    back = L[-1][1]
    relabel = back < EPS
    fixed = L[:-1] if relabel else L
    for gi, v in fixed:
        if gi == g:
            return v
    return back if relabel else 0.0


def put(t, g, v):
    if t and t[-1][1] < EPS:
        t[-1] = [g, v]
    else:
        t.append([g, v])


class Grid:
    def __init__(self, Lx, Ly, Lz):
        self.L = (Lx, Ly, Lz)
        self.l = (Lx + 6, Ly + 6, Lz + 6)

    def interior(self, x, y, z):
        return all(3 <= p < q - 3 for p, q in zip((x, y, z), self.l))

    def voxels(self):
        Lx, Ly, Lz = self.L
        for x in range(3, Lx + 3):
            for y in range(3, Ly + 3):
                for z in range(3, Lz + 3):
                    yield (x, y, z)

    def cells(self):
        lx, ly, lz = self.l
        for x in range(lx):
            for y in range(ly):
                for z in range(lz):
                    yield (x, y, z)

    def image(self, c):
        return tuple(p + L if p < 3 else p - L if p >= L + 3 else p for p, L in zip(c, self.L))

    def d(self, c):
        Lx, Ly, Lz = self.L
        return ((c[0] - 3) * Ly + c[1] - 3) * Lz + c[2] - 3


def halo(G, A):
    for c in G.cells():
        if G.interior(*c):
            continue
        t = []
        for g, v in A[G.image(c)]:
            put(t, g, v)
        A[c] = t


def substep(G, A, defeng, gid, rho_avg, P):
    """One Update_grid + boundary_conditions + swap; returns the new lists and
    changes defeng (interior resets) and rho_avg in place."""
    M_coeff, kappa, m, E_disl = coeffs(P)
    init_gid = P["init_gid"]
    vox = list(G.voxels())
    # candidate grains (fact 1)
    Gl = {}
    for v in vox:
        lst = []
        for i in range(v[0] - 1, v[0] + 2):
            for j in range(v[1] - 1, v[1] + 2):
                for k in range(v[2] - 1, v[2] + 2):
                    for g, val in A[(i, j, k)]:
                        if val > EPS and g not in lst:
                            lst.append(g)
        Gl[v] = lst
    # update branch: least fixed point (fact 3)
    act = {v: len(A[v]) >= 2 for v in vox}
    while True:
        changed = False
        for v in vox:
            L0 = A[v]
            if act[v] or len(L0) != 1 or not (L0[0][1] >= EPS):
                continue
            own = L0[0][0]
            for e in EARLY:
                u = (v[0] + e[0], v[1] + e[1], v[2] + e[2])
                if G.interior(*u) and act[u] and any(g != own for g in Gl[u]):
                    act[v] = True
                    changed = True
                    break
        if not changed:
            break
    # the list at each voxel's turn
    turn = {}
    for v in vox:
        L0 = A[v]
        relabel = bool(L0) and L0[-1][1] < EPS
        fixed = [g for g, _ in (L0[:-1] if relabel else L0)]
        found = None
        for e in EARLY:
            u = (v[0] + e[0], v[1] + e[1], v[2] + e[2])
            if not (G.interior(*u) and act[u]):
                continue
            for g in reversed(Gl[u]):
                if g not in fixed:
                    found = g
                    break
            if found is not None:
                break
        ent = [list(p) for p in L0]
        if relabel:
            if found is not None:
                ent[-1][0] = found
        elif found is not None:
            ent.append([found, 0.0])
        turn[v] = ent
    # rho_avg writers (fact 4)
    writers = []
    for v in vox:
        if len(turn[v]) > 1:
            for k, (g, _) in enumerate(turn[v]):
                writers.append(((g - 1, G.d(v)), (v, k)))
    writers.sort()
    wkeys = [w[0] for w in writers]
    trig = set()
    while True:
        avg = {}
        for v in vox:
            ent = turn[v]
            if len(ent) <= 1:
                continue
            a = [0.0] * len(ent)
            cn = [0] * len(ent)
            for i in range(v[0] - 3, v[0] + 4):
                for j in range(v[1] - 3, v[1] + 4):
                    for k in range(v[2] - 3, v[2] + 4):
                        p = (i, j, k)
                        e = defeng[p]
                        if G.interior(*p) and p in trig and G.d(p) < G.d(v):
                            e = 100.0
                        for q, (g, _) in enumerate(ent):
                            if gid[p] == g:
                                a[q] += e
                                cn[q] += 1
            avg[v] = (a, cn)

        def wval(ref):
            w, k = ref
            a, cn = avg[w]
            return a[k] / cn[k] if cn[k] > 0 else 0.0

        def rho_at(g, v):
            i = bisect.bisect_right(wkeys, (g - 1, G.d(v)))
            if i > 0 and wkeys[i - 1][0] == g - 1:
                return wval(writers[i - 1][1])
            return rho_avg[g - 1]

        B = {}
        new_trig = set()
        for v in vox:
            ent = turn[v]
            t = []
            if len(ent) > 1:
                a, cn = avg[v]
                sum_rex = sum1 = h_phi_def = 0.0
                for q, (g, val) in enumerate(ent):
                    if cn[q] > 0 and a[q] > 0.0:
                        sum_rex += val * val
                        sum1 += val * val * val * (10.0 - 15.0 * val + 6.0 * val * val)
                        h_phi_def += (val * val * val * (10.0 - 15.0 * val + 6.0 * val * val)) * (a[q] / cn[q])
                count = sum(1 for g in Gl[v] if g > init_gid)
                for g in Gl[v]:
                    s = lookupv(A[v], g)
                    h_phi = s * s * s * (10.0 - 15.0 * s + 6.0 * s * s)
                    dh_phi = s * s * (30.0 - 60.0 * s + 30.0 * s * s)
                    args = [lookupv(A[(v[0] + o[0], v[1] + o[1], v[2] + o[2])], g) for o in LAPN] + [s]
                    lap = laplacian(args)
                    ra = rho_at(g, v)
                    if ra > 0.0:
                        c = s
                        value = c - M_coeff * (m * (c * c * c - c + 3.0 * c * (sum_rex - c * c)) - kappa * lap +
                                               E_disl * dh_phi * ((sum1 - h_phi) * ra - (h_phi_def - h_phi * ra)) /
                                               (sum1 * sum1))
                        if count > 0 and g > init_gid and c < 0.5 and value > 0.5:
                            new_trig.add(v)
                        if value > 1e-4:
                            put(t, g, value)
            elif len(ent) == 1:
                for g in Gl[v]:
                    put(t, g, lookupv(A[v], g))
            B[v] = t
        if new_trig == trig:
            break
        trig = new_trig
    for v in trig:
        defeng[v] = 100.0
    last = {}
    for key, ref in writers:
        last[key[0]] = ref
    for g0, ref in last.items():
        rho_avg[g0] = wval(ref)
    for c in G.cells():
        if not G.interior(*c):
            B[c] = []
    halo(G, B)
    return B


def mesh(G, arr):
    for c in G.cells():
        if not G.interior(*c):
            arr[c] = arr[G.image(c)]


def run(gid_field, rho, steps, P, rho_avg):
    """PF_XRX from a grain field: returns (gid_new, lists, defeng, rho_avg)."""
    Lx, Ly, Lz = gid_field.shape
    G = Grid(Lx, Ly, Lz)
    A, defeng, gid = {}, {}, {}
    for c in G.cells():
        A[c] = []
        defeng[c] = 0.0
        gid[c] = 0
    for v in G.voxels():
        x, y, z = (p - 3 for p in v)
        A[v] = [[int(gid_field[x, y, z]), 1.0]]
        defeng[v] = float(rho[x, y, z])
        gid[v] = int(gid_field[x, y, z])
    halo(G, A)
    mesh(G, defeng)
    mesh(G, gid)
    ra = [float(r) for r in rho_avg]
    gid_new = gid_field.copy()
    for _ in range(steps):
        A = substep(G, A, defeng, gid, ra, P)
        for v in G.voxels():
            best, mx = 0, 0.0
            for g, val in A[v]:
                if mx < val:
                    mx, best = val, g
            gid_new[v[0] - 3, v[1] - 3, v[2] - 3] = best
        mesh(G, defeng)
    return gid_new, A, defeng, ra


def run_lists(n, idx, val, GID, defeng_in, steps, P, rho_avg):
    """The substeps from given interior lists (the oracle's orc_pf_run_lists)."""
    Lx, Ly, Lz = GID.shape
    G = Grid(Lx, Ly, Lz)
    A, defeng, gid = {}, {}, {}
    for c in G.cells():
        A[c] = []
        defeng[c] = 0.0
        gid[c] = 0
    for v in G.voxels():
        x, y, z = (p - 3 for p in v)
        A[v] = [[int(idx[x, y, z, k]), float(val[x, y, z, k])] for k in range(int(n[x, y, z]))]
        defeng[v] = float(defeng_in[x, y, z])
        gid[v] = int(GID[x, y, z])
    halo(G, A)
    mesh(G, defeng)
    mesh(G, gid)
    ra = [float(r) for r in rho_avg]
    gid_new = GID.copy()
    for _ in range(steps):
        A = substep(G, A, defeng, gid, ra, P)
        for v in G.voxels():
            best, mx = 0, 0.0
            for g, x_ in A[v]:
                if mx < x_:
                    mx, best = x_, g
            gid_new[v[0] - 3, v[1] - 3, v[2] - 3] = best
        mesh(G, defeng)
    return gid_new, A, defeng, ra
