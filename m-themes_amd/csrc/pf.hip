// pf.hip -- the sparse phase-field recrystallisation step PF_XRX
// (src/pf_xrx.c:1469-1698) on one GPU, bit for bit the one-rank reference.
//
// The reference keeps each voxel's order parameters in a PFdata list
// (src/PFgrid.h:54-94) and updates them in one x-major sweep, Update_grid
// (src/pf_xrx.c:895-1311), whose neighbour reads go through PFdata's
// non-const lookup: a grain absent from a list re-labels the list's last
// entry when that entry's value is below 5e-3, else a zero entry is
// appended.  The sweep is sequential as written; it is computed here in
// parallel from four facts about it:
//   1. Lookups never create or re-label an entry above 5e-3, so the set of
//      grains above 5e-3 around a voxel (its candidate grains G) and every
//      list's entries above 5e-3 are those of the start of the sweep.
//   2. The value a lookup of grain g returns from a list L depends on L's
//      start state only: L's value of g when g is an entry other than a
//      last entry below 5e-3; otherwise that last entry's value when it is
//      below 5e-3, else 0.
//   3. A voxel takes the update branch when its list holds more than one
//      entry at its turn.  A one-entry list above 5e-3 gains a zero entry at
//      the first lookup of another grain, and only voxels on the update
//      branch look into their neighbours, so the branch is the least fixed
//      point of "an earlier update-branch neighbour has a candidate grain I
//      lack" -- a monotone iteration over the sweep order.  The list's
//      re-labelled or appended last entry at that turn carries the last
//      such grain looked up.
//   4. rho_avg[g-1] at a voxel's turn was written by the last voxel before
//      it (or itself) whose list held g: sorted (grain, position) keys and a
//      binary search.  The stored-energy reset to 100 (a new grain passing
//      0.5) feeds later voxels' 7x7x7 averages; the triggers are iterated to
//      their fixed point over the sweep order.
// Lists are structure of arrays, entry k of cell c at k*NC + c, over the
// reference's padded grid (three halo planes each side).  All arithmetic is
// IEEE double in the reference's operation order: this file is built with
// -ffp-contract=off.
#include <hipcub/hipcub.hpp>
#include <limits.h>

#include "evpgpu_impl.h"

namespace {

constexpr double PF_EPS = 5.0E-3; // src/PFgrid.h:50
constexpr int CAP = EVPGPU_PF_CAP;
constexpr int NT = 256;
constexpr int MAX_PASSES = 4096;

enum { PFE_CAP = 1, PFE_GRAIN = 2, PFE_SUM = 4 };

struct Dims {
    int Lx, Ly, Lz, lx, ly, lz;
    long long NC, NV;
};

struct Lists {
    int *cnt;
    int *idx;
    double *val;
};

__host__ __device__ inline long long cix(const Dims &g, int x, int y, int z)
{
    return ((long long)x * g.ly + y) * g.lz + z;
}

// interior voxel d -> padded coordinates
__device__ inline void vox(const Dims &g, long long d, int &x, int &y, int &z)
{
    z = (int)(d % g.Lz) + 3;
    const long long r = d / g.Lz;
    y = (int)(r % g.Ly) + 3;
    x = (int)(r / g.Ly) + 3;
}

__device__ inline bool interior(const Dims &g, int x, int y, int z)
{
    return x >= 3 && x < g.lx - 3 && y >= 3 && y < g.ly - 3 && z >= 3 && z < g.lz - 3;
}

__device__ inline long long vid(const Dims &g, int x, int y, int z)
{
    return ((long long)(x - 3) * g.Ly + (y - 3)) * g.Lz + (z - 3);
}

// A list being built by assignments list[gr] = v of distinct grains (the
// temp lists of Update_grid, the halo copies of boundary_conditions),
// written straight into the structure-of-arrays lists at cell c: an
// assignment only ever changes the last entry (a re-label when it is below
// EPSILON) or appends one, so every entry but the last is final and only
// the last is held in registers.
struct Build {
    int *cnt, *gidx;
    double *gval;
    long long NC, c;
    int n, last_idx;
    double last_val;
    __device__ Build(const Lists &L, long long NC_, long long c_)
        : cnt(L.cnt), gidx(L.idx), gval(L.val), NC(NC_), c(c_), n(0), last_idx(0), last_val(0.0)
    {
    }
    __device__ bool put(int gr, double v)
    {
        if (n > 0 && last_val < PF_EPS) {
            last_idx = gr;
            last_val = v;
            return true;
        }
        if (n >= CAP) return false;
        if (n > 0) {
            gidx[(long long)(n - 1) * NC + c] = last_idx;
            gval[(long long)(n - 1) * NC + c] = last_val;
        }
        last_idx = gr;
        last_val = v;
        n++;
        return true;
    }
    __device__ void finish()
    {
        if (n > 0) {
            gidx[(long long)(n - 1) * NC + c] = last_idx;
            gval[(long long)(n - 1) * NC + c] = last_val;
        }
        cnt[c] = n;
    }
};

// the 9 face/edge neighbours that come before a voxel in the x-major sweep,
// closest first (linear offsets -1, -Lz+1, -Lz, -Lz-1, -LyLz+Lz, -LyLz+1,
// -LyLz, -LyLz-1, -LyLz-Lz for lengths >= 3)
__constant__ int EARLY[9][3] = {{0, 0, -1},  {0, -1, 1}, {0, -1, 0}, {0, -1, -1}, {-1, 1, 0},
                                {-1, 0, 1},  {-1, 0, 0}, {-1, 0, -1}, {-1, -1, 0}};

// does neighbour u (interior voxel) look up a grain outside {own}?
__device__ inline bool has_other(const Dims &g, const int *gcnt, const int *gidx, long long u, int own)
{
    const int n = gcnt[u];
    for (int k = 0; k < n; k++)
        if (gidx[(long long)k * g.NV + u] != own) return true;
    return false;
}

// Start of the update-branch fixed point: the voxels whose list holds two
// entries or more, and for each one-entry list above EPSILON the mask of its
// nine earlier neighbours (EARLY order) whose candidate grains include
// another grain -- the neighbours whose joining makes it join.
__global__ __launch_bounds__(NT) void k_act_init(Dims g, Lists A, const int *gcnt, const int *gidx, int *act,
                                                 int *emask)
{
    for (long long d = blockIdx.x * (long long)NT + threadIdx.x; d < g.NV; d += (long long)gridDim.x * NT) {
        int x, y, z;
        vox(g, d, x, y, z);
        const long long c = cix(g, x, y, z);
        const int n = A.cnt[c];
        act[d] = n >= 2;
        int mask = 0;
        if (n == 1 && A.val[c] >= PF_EPS) {
            const int own = A.idx[c];
            for (int e = 0; e < 9; e++) {
                const int ux = x + EARLY[e][0], uy = y + EARLY[e][1], uz = z + EARLY[e][2];
                if (interior(g, ux, uy, uz) && has_other(g, gcnt, gidx, vid(g, ux, uy, uz), own)) mask |= 1 << e;
            }
        }
        emask[d] = mask;
    }
}

// One pass of the update-branch fixed point: a one-entry list above
// EPSILON joins when an earlier update-branch neighbour looks up another
// grain.  A workgroup of 16 waves owns a 4 x 4 tile of z rows, one wave per
// row in chunks of 64 voxels: the eight earlier neighbours in other rows give
// each voxel a "generate" bit, the row predecessor a "propagate" bit, and the
// chain along the row (a_z = gen_z | prop_z a_(z-1)) is resolved by a
// doubling scan over the wave's ballots; the workgroup repeats this until its
// tile stops changing, so only chains that leave a tile take further passes.
constexpr int ACT_NT = 1024, ACT_T = 4;

__global__ __launch_bounds__(ACT_NT) void k_act_iter(Dims g, const int *emask, int *act, int *changed)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ntx = (g.Lx + ACT_T - 1) / ACT_T;
    const int bx = blockIdx.x % ntx, by = blockIdx.x / ntx;
    const int x = 3 + bx * ACT_T + w / ACT_T, y = 3 + by * ACT_T + w % ACT_T;
    const bool row_ok = x < g.lx - 3 && y < g.ly - 3; // wave-uniform
    bool any = false;
    for (;;) {
        bool moved = false;
        if (row_ok) {
            unsigned long long carry = 0;
            for (int z0 = 0; z0 < g.Lz; z0 += 64) {
                const int z = z0 + lane + 3;
                const bool in = z0 + lane < g.Lz;
                bool cur = false, gen = false, prop = false;
                long long d = 0;
                if (in) {
                    d = vid(g, x, y, z);
                    cur = __atomic_load_n(&act[d], __ATOMIC_RELAXED) != 0;
                    gen = cur;
                    const int mask = emask[d];
                    if (!cur && mask) {
                        for (int e = 1; e < 9 && !gen; e++)
                            if ((mask >> e) & 1)
                                gen = __atomic_load_n(&act[vid(g, x + EARLY[e][0], y + EARLY[e][1], z + EARLY[e][2])],
                                                      __ATOMIC_RELAXED) != 0;
                        prop = !gen && (mask & 1);
                    }
                }
                unsigned long long G = __ballot(gen), P = __ballot(prop);
#pragma unroll
                for (int sft = 1; sft < 64; sft <<= 1) {
                    G |= P & (G << sft);
                    P &= (P << sft) | ((1ull << sft) - 1ull);
                }
                const unsigned long long res = G | (carry ? P : 0ull);
                if (in && !cur && ((res >> lane) & 1ull)) {
                    __atomic_store_n(&act[d], 1, __ATOMIC_RELAXED);
                    moved = true;
                }
                const int last = g.Lz - 1 - z0 < 63 ? g.Lz - 1 - z0 : 63;
                carry = (res >> last) & 1ull;
            }
        }
        any = any || moved;
        if (!__syncthreads_or(moved)) break;
    }
    if (any) *changed = 1;
}

// The list at the voxel's turn: its start entries, with the last entry
// re-labelled (start value below EPSILON) or a zero entry appended, to the
// last grain outside the fixed entries that an earlier update-branch
// neighbour looked up.  nn: entries; spos: position of that entry or -1;
// sg: its grain; wn: rho_avg writes (nn on the update branch).
__global__ __launch_bounds__(NT) void k_slot(Dims g, Lists A, const int *gcnt, const int *gidx, const int *act, int *nn, int *spos,
                       int *sg, int *wn, int *nupd, int *err)
{
    // The loads go out in two batches instead of one chain per neighbour:
    // first the voxel's own list (its first OWN entries whatever the count:
    // the lists hold CAP slots per cell) and the 9 early neighbours' activity
    // flags and candidate counts, then each active neighbour's last candidate
    // (the first one the search tries).
    constexpr int OWN = 4;
    for (long long d = blockIdx.x * (long long)NT + threadIdx.x; d < g.NV; d += (long long)gridDim.x * NT) {
        int x, y, z;
        vox(g, d, x, y, z);
        const long long c = cix(g, x, y, z);
        const int n0 = A.cnt[c];
        int own[OWN];
        double ownv[OWN];
#pragma unroll
        for (int f = 0; f < OWN; f++) {
            own[f] = A.idx[(long long)f * g.NC + c];
            ownv[f] = A.val[(long long)f * g.NC + c];
        }
        long long u9[9];
        bool in9[9];
        int a9[9], n9[9];
#pragma unroll
        for (int e = 0; e < 9; e++) {
            const int ux = x + EARLY[e][0], uy = y + EARLY[e][1], uz = z + EARLY[e][2];
            in9[e] = interior(g, ux, uy, uz);
            u9[e] = in9[e] ? vid(g, ux, uy, uz) : d;
            a9[e] = act[u9[e]];
            n9[e] = gcnt[u9[e]];
        }
        int top9[9];
#pragma unroll
        for (int e = 0; e < 9; e++) {
            top9[e] = 0;
            if (in9[e] && a9[e] && n9[e] > 0) top9[e] = gidx[(long long)(n9[e] - 1) * g.NV + u9[e]];
        }
        double lastv = 1.0;
        int lasti = 0;
        if (n0 > 0) {
#pragma unroll
            for (int f = 0; f < OWN; f++)
                if (f == n0 - 1) {
                    lastv = ownv[f];
                    lasti = own[f];
                }
            if (n0 > OWN) {
                lastv = A.val[(long long)(n0 - 1) * g.NC + c];
                lasti = A.idx[(long long)(n0 - 1) * g.NC + c];
            }
        }
        const bool relabel = n0 > 0 && lastv < PF_EPS;
        const int nf = relabel ? n0 - 1 : n0;
        // is grain gr one of the first nf entries of the voxel's list?
        auto fixed_grain = [&](int gr) {
            bool fixed = false;
#pragma unroll
            for (int f = 0; f < OWN; f++) fixed = fixed || (f < nf && own[f] == gr);
            for (int f = OWN; f < nf; f++) fixed = fixed || A.idx[(long long)f * g.NC + c] == gr;
            return fixed;
        };
        int found = 0;
#pragma unroll
        for (int e = 0; e < 9; e++) {
            if (found || !in9[e] || !a9[e]) continue;
            const long long u = u9[e];
            const int n = n9[e];
            for (int k = n - 1; k >= 0; k--) {
                const int gr = k == n - 1 ? top9[e] : gidx[(long long)k * g.NV + u];
                if (!fixed_grain(gr)) {
                    found = gr;
                    break;
                }
            }
        }
        int n = n0, p = -1, s = 0;
        if (relabel) {
            p = n0 - 1;
            s = found ? found : lasti;
        } else if (found) {
            if (n0 == CAP) {
                atomicOr(err, PFE_CAP); // the appended entry has no room
            } else {
                p = n0;
                s = found;
                n = n0 + 1;
            }
        }
        nn[d] = n;
        spos[d] = p;
        sg[d] = s;
        wn[d] = n > 1 ? n : 0;
        if (n > 1) atomicAdd(nupd, 1);
    }
}

// entry k of the list at the voxel's turn
__device__ inline void entry(const Dims &g, const Lists &A, long long c, int k, int p, int s, int &gr, double &v)
{
    if (k == p) {
        gr = s;
        v = k < A.cnt[c] ? A.val[(long long)k * g.NC + c] : 0.0;
    } else {
        gr = A.idx[(long long)k * g.NC + c];
        v = A.val[(long long)k * g.NC + c];
    }
}

// rho_avg writes as (grain - 1, sweep position) keys
__global__ __launch_bounds__(NT) void k_emit(Dims g, Lists A, const int *nn, const int *spos, const int *sg, const int *wn,
                       const long long *woff, unsigned long long *keys, int *refs, int n_rho, int *err)
{
    for (long long d = blockIdx.x * (long long)NT + threadIdx.x; d < g.NV; d += (long long)gridDim.x * NT) {
        const int n = wn[d];
        if (!n) continue;
        int x, y, z;
        vox(g, d, x, y, z);
        const long long c = cix(g, x, y, z);
        long long o = woff[d];
        for (int k = 0; k < n; k++) {
            int gr;
            double v;
            entry(g, A, c, k, spos[d], sg[d], gr, v);
            if (gr < 1 || gr > n_rho) {
                atomicOr(err, PFE_GRAIN);
                gr = 1;
            }
            keys[o] = (unsigned long long)(gr - 1) * (unsigned long long)g.NV + (unsigned long long)d;
            refs[o] = (int)(d * CAP + k);
            o++;
        }
    }
}

// Avg and count_dd of each list entry over the 7x7x7 box (src/pf_xrx.c:1042-1077).
// A workgroup takes a 4 x 8 x 8 tile of voxels and stages the grain field,
// the stored energy and the reset flags of the tile's box (10 x 14 x 14 cells)
// in LDS.  The stored energy a voxel reads at a cell is 100 where an earlier
// voxel of this sweep reset it (trig), else the buffer's.  With tprev (the
// resets the previous pass used) a tile whose box saw no change keeps its
// averages.
constexpr int TX = 4, TY = 8, TZ = 8, HX = TX + 6, HY = TY + 6, HZ = TZ + 6, HC = HX * HY * HZ;
static_assert(TX * TY * TZ == NT, "one voxel per thread");

// The lists of a 4 x 8 x 8 voxel tile and of its face/edge neighbours (a
// 6 x 10 x 10 cell box) staged in LDS: counts and the first LE entries;
// entries beyond LE (rare) are read from memory.  31 KB.
constexpr int LX = TX + 2, LY = TY + 2, LZ = TZ + 2, LC = LX * LY * LZ, LE = 4;
struct ListTile {
    int cnt[LC];
    int idx[LE * LC];
    double val[LE * LC];
};

// tile b's first voxel in padded coordinates
__device__ inline void tile_origin(const Dims &g, int b, int &x0, int &y0, int &z0)
{
    const int ntx = (g.Lx + TX - 1) / TX, nty = (g.Ly + TY - 1) / TY;
    x0 = 3 + (b % ntx) * TX;
    y0 = 3 + ((b / ntx) % nty) * TY;
    z0 = 3 + (b / (ntx * nty)) * TZ;
}

__device__ inline void load_lists(ListTile &T, const Dims &g, const Lists &A, int x0, int y0, int z0)
{
    for (int q = threadIdx.x; q < LC; q += NT) {
        const int hz = q % LZ, hy = (q / LZ) % LY, hx = q / (LZ * LY);
        const int i = x0 - 1 + hx, j = y0 - 1 + hy, m = z0 - 1 + hz;
        int n = 0;
        if (i < g.lx && j < g.ly && m < g.lz) {
            const long long c = cix(g, i, j, m);
            n = A.cnt[c];
#pragma unroll
            for (int k = 0; k < LE; k++)
                if (k < n) {
                    T.idx[k * LC + q] = A.idx[(long long)k * g.NC + c];
                    T.val[k * LC + q] = A.val[(long long)k * g.NC + c];
                }
        }
        T.cnt[q] = n;
    }
}

// tile cell of the voxel at tile offset (tx, ty, tz) plus (dx, dy, dz)
__device__ inline int tq(int tx, int ty, int tz, int dx, int dy, int dz)
{
    return ((tx + 1 + dx) * LY + (ty + 1 + dy)) * LZ + (tz + 1 + dz);
}

__device__ inline int tile_idx(const ListTile &T, const Dims &g, const Lists &A, int q, long long c, int k)
{
    return k < LE ? T.idx[k * LC + q] : A.idx[(long long)k * g.NC + c];
}

__device__ inline double tile_val(const ListTile &T, const Dims &g, const Lists &A, int q, long long c, int k)
{
    return k < LE ? T.val[k * LC + q] : A.val[(long long)k * g.NC + c];
}

// the value a PFdata lookup of grain gr returns from a start-of-sweep list
// (tile cell q, memory cell c): its value when gr is an entry other than a
// last entry below EPSILON, else that last entry's value when below, else 0
__device__ inline double lookup_t(const ListTile &T, const Dims &g, const Lists &A, int q, long long c, int gr)
{
    const int n = T.cnt[q];
    if (n == 0) return 0.0;
    const double back = tile_val(T, g, A, q, c, n - 1);
    const bool relabel = back < PF_EPS;
    const int nf = relabel ? n - 1 : n;
    for (int k = 0; k < nf; k++)
        if (tile_idx(T, g, A, q, c, k) == gr) return tile_val(T, g, A, q, c, k);
    return relabel ? back : 0.0;
}

// candidate grains: values above EPSILON in the 3x3x3 neighbourhood, in
// order of first appearance (src/pf_xrx.c:940-975); lists from the tile
__global__ __launch_bounds__(NT) void k_glist(Dims g, Lists A, int *gcnt, int *gidx, int *err)
{
    __shared__ ListTile T;
    int x0, y0, z0;
    tile_origin(g, blockIdx.x, x0, y0, z0);
    load_lists(T, g, A, x0, y0, z0);
    __syncthreads();
    const int tz = threadIdx.x % TZ, ty = (threadIdx.x / TZ) % TY, tx = threadIdx.x / (TZ * TY);
    const int x = x0 + tx, y = y0 + ty, z = z0 + tz;
    if (!interior(g, x, y, z)) return;
    const long long d = vid(g, x, y, z);
    int G[CAP];
    int nG = 0;
    bool over = false;
    for (int i = -1; i <= 1; ++i)
        for (int j = -1; j <= 1; ++j)
            for (int m = -1; m <= 1; ++m) {
                const int q = tq(tx, ty, tz, i, j, m);
                const long long c = cix(g, x + i, y + j, z + m);
                const int n = T.cnt[q];
                for (int k = 0; k < n; ++k)
                    if (tile_val(T, g, A, q, c, k) > PF_EPS) {
                        const int index = tile_idx(T, g, A, q, c, k);
                        bool seen = false;
#pragma unroll
                        for (int it = 0; it < CAP; it++) seen = seen || (it < nG && G[it] == index);
                        if (!seen) {
                            if (nG == CAP) {
                                over = true;
                            } else {
#pragma unroll
                                for (int it = 0; it < CAP; it++)
                                    if (it == nG) G[it] = index;
                                nG++;
                            }
                        }
                    }
            }
    if (over) atomicOr(err, PFE_CAP);
    gcnt[d] = nG;
#pragma unroll
    for (int k = 0; k < CAP; k++)
        if (k < nG) gidx[(long long)k * g.NV + d] = G[k];
}


// the sums of one voxel over its box, for lists of at most M entries (the
// entries' grains, sums and counts stay in registers)
template <int M>
__device__ inline void avg_box(const Dims &g, const Lists &A, long long c, long long d, int n, int p, int s, int tx,
                               int ty, int tz, int x, int y, int z, const int *s_gid, const double *s_de,
                               const unsigned char *s_tf, double *avg, int *cdd)
{
    int gk[M];
    double a[M];
    int cn[M];
#pragma unroll
    for (int k = 0; k < M; k++) {
        gk[k] = INT_MIN; // matches no grain, nor the -1 of cells outside the grid
        if (k < n) {
            double v;
            entry(g, A, c, k, p, s, gk[k], v);
        }
        a[k] = 0.0;
        cn[k] = 0;
    }
    if constexpr (M > 4) {
        // long lists (rare): the cell-by-cell loop, which holds fewer registers
        for (int i = 0; i < 7; ++i)
            for (int j = 0; j < 7; ++j)
                for (int m = 0; m < 7; ++m) {
                    const int q = ((tx + i) * HY + (ty + j)) * HZ + (tz + m);
                    const int gp = s_gid[q];
                    bool hit = false;
#pragma unroll
                    for (int k = 0; k < M; k++) hit = hit || gp == gk[k];
                    if (!hit) continue;
                    double e = s_de[q];
                    if (s_tf[q] && vid(g, x - 3 + i, y - 3 + j, z - 3 + m) < d) e = 100.0;
#pragma unroll
                    for (int k = 0; k < M; k++)
                        if (gp == gk[k]) {
                            a[k] += e;
                            cn[k] += 1;
                        }
                }
    } else {
        for (int i = 0; i < 7; ++i)
            for (int j = 0; j < 7; ++j) {
                // one z row of the box: its 7 cells' LDS reads go out together,
                // then the sums in the cells' order (the reads are unconditional,
                // the sums predicated)
                const int q0 = ((tx + i) * HY + (ty + j)) * HZ + tz;
                int gp[7];
                double ev[7];
                unsigned char tf[7];
#pragma unroll
                for (int m = 0; m < 7; ++m) {
                    gp[m] = s_gid[q0 + m];
                    ev[m] = s_de[q0 + m];
                    tf[m] = s_tf[q0 + m];
                }
#pragma unroll
                for (int m = 0; m < 7; ++m) {
                    double e = ev[m];
                    if (tf[m] && vid(g, x - 3 + i, y - 3 + j, z - 3 + m) < d) e = 100.0;
#pragma unroll
                    for (int k = 0; k < M; k++)
                        if (gp[m] == gk[k]) {
                            a[k] += e;
                            cn[k] += 1;
                        }
                }
            }
    }
#pragma unroll
    for (int k = 0; k < M; k++)
        if (k < n) {
            avg[(long long)k * g.NV + d] = a[k];
            cdd[(long long)k * g.NV + d] = cn[k];
        }
}

__global__ __launch_bounds__(NT) void k_avg(Dims g, Lists A, const int *wn, const int *spos, const int *sg,
                                            const int *gid, const double *defeng, const int *trig,
                                            const int *tprev, double *avg, int *cdd)
{
    __shared__ int s_gid[HC];
    __shared__ double s_de[HC];
    __shared__ unsigned char s_tf[HC];
    const int ntx = (g.Lx + TX - 1) / TX, nty = (g.Ly + TY - 1) / TY;
    const int b = blockIdx.x;
    const int bx = b % ntx, by = (b / ntx) % nty, bz = b / (ntx * nty);
    // padded coordinates of the tile's first voxel and of its box
    const int x0 = 3 + bx * TX, y0 = 3 + by * TY, z0 = 3 + bz * TZ;
    int changed = tprev ? 0 : 1;
    for (int q = threadIdx.x; q < HC; q += NT) {
        const int hz = q % HZ, hy = (q / HZ) % HY, hx = q / (HZ * HY);
        const int i = x0 - 3 + hx, j = y0 - 3 + hy, m = z0 - 3 + hz;
        int gg = -1, tf = 0;
        double e = 0.0;
        if (i < g.lx && j < g.ly && m < g.lz) {
            const long long c = cix(g, i, j, m);
            gg = gid[c];
            e = defeng[c];
            if (interior(g, i, j, m)) {
                const long long v = vid(g, i, j, m);
                tf = trig[v];
                if (tprev && tprev[v] != tf) changed = 1;
            }
        }
        s_gid[q] = gg;
        s_de[q] = e;
        s_tf[q] = (unsigned char)tf;
    }
    if (!__syncthreads_or(changed)) return;
    const int tz = threadIdx.x % TZ, ty = (threadIdx.x / TZ) % TY, tx = threadIdx.x / (TZ * TY);
    const int x = x0 + tx, y = y0 + ty, z = z0 + tz;
    if (!interior(g, x, y, z)) return;
    const long long d = vid(g, x, y, z);
    const int n = wn[d];
    if (!n) return;
    const long long c = cix(g, x, y, z);
    if (n <= 4) avg_box<4>(g, A, c, d, n, spos[d], sg[d], tx, ty, tz, x, y, z, s_gid, s_de, s_tf, avg, cdd);
    else avg_box<CAP>(g, A, c, d, n, spos[d], sg[d], tx, ty, tz, x, y, z, s_gid, s_de, s_tf, avg, cdd);
}

__device__ inline double writer_value(const Dims &g, const double *avg, const int *cdd, int ref)
{
    const long long w = ref / CAP;
    const int k = ref % CAP;
    const int cn = cdd[(long long)k * g.NV + w];
    return cn > 0 ? avg[(long long)k * g.NV + w] / cn : 0.0;
}

// rho_avg[gr-1] at voxel d's turn: the last writer of grain gr at or before
// d, searched in the grain's own run [gfirst, gend) of the sorted keys (a
// search over all keys takes ~21 dependent loads at 128^3, the grain's run
// ~13)
__device__ inline double rho_at(const Dims &g, const unsigned long long *keys, const long long *gfirst,
                                const long long *gend, const int *refs, const double *avg, const int *cdd,
                                const double *rho_in, int gr, long long d)
{
    const unsigned long long key = (unsigned long long)(gr - 1) * (unsigned long long)g.NV + (unsigned long long)d;
    const long long lo0 = gfirst[gr - 1];
    long long lo = lo0, hi = gend[gr - 1]; // first index of the run with keys[i] > key
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (keys[mid] <= key) lo = mid + 1;
        else hi = mid;
    }
    if (lo > lo0) return writer_value(g, avg, cdd, refs[lo - 1]);
    return rho_in[gr - 1];
}

// each grain's run [gfirst, gend) in the sorted writer keys (both cleared to
// 0 first: an empty run for grains without writers)
__global__ __launch_bounds__(NT) void k_grange(const unsigned long long *keys, long long nw, long long NV,
                                               long long *gfirst, long long *gend)
{
    for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < nw; i += (long long)gridDim.x * NT) {
        const unsigned long long gr = keys[i] / (unsigned long long)NV;
        if (i == 0 || keys[i - 1] / (unsigned long long)NV != gr) gfirst[gr] = i;
        if (i + 1 == nw || keys[i + 1] / (unsigned long long)NV != gr) gend[gr] = i + 1;
    }
}


struct Coef {
    double M_coeff, kappa, m, E_disl;
    int init_gid, n_rho;
};

// The cells of the laplacian call (src/pf_xrx.c:1150-1160) in the order its
// body sums them: the faces right, left, front, back, up, down; then the
// edges in the order laplacian() adds its parameters, whose names 7-18 do
// not match the arguments: center_up_front, center_down_front,
// center_up_back, center_down_back, left_down, left_up, left_front,
// left_back, right_up, right_back, right_down, right_front.
__constant__ int LAPQ[18][3] = {{1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0}, {0, 0, 1},   {0, 0, -1},
                                {0, -1, 1},  {0, -1, -1}, {0, 1, 1},   {0, 1, -1}, {-1, 0, -1}, {-1, 0, 1},
                                {-1, -1, 0}, {-1, 1, 0},  {1, 0, 1},   {1, 1, 0},  {1, 0, -1},  {1, -1, 0}};

// the voxel update of src/pf_xrx.c:1034-1246 into the temp lists B
__global__ __launch_bounds__(NT) void k_update(Dims g, Lists A, Lists B, Coef P, const int *gcnt, const int *gidx, const int *nn,
                         const int *spos, const int *sg, const double *avg, const int *cdd,
                         const unsigned long long *keys, const long long *gfirst, const long long *gend,
                         const int *refs, const double *rho_in, int *trig_out, int *err)
{
    __shared__ ListTile T;
    int x0, y0, z0;
    tile_origin(g, blockIdx.x, x0, y0, z0);
    load_lists(T, g, A, x0, y0, z0);
    __syncthreads();
    const int tz = threadIdx.x % TZ, ty = (threadIdx.x / TZ) % TY, tx = threadIdx.x / (TZ * TY);
    const int x = x0 + tx, y = y0 + ty, z = z0 + tz;
    if (interior(g, x, y, z)) {
        const long long d = vid(g, x, y, z);
        const long long c = cix(g, x, y, z);
        const int q0 = tq(tx, ty, tz, 0, 0, 0);
        const int NN = nn[d];
        const int nG = gcnt[d];
        Build t(B, g.NC, c);
        bool ok = true;
        int trig = 0;
        if (NN > 1) {
            double sum_rex = 0.0, sum1 = 0.0, h_phi_def = 0.0;
            for (int k = 0; k < NN; k++) {
                const int cn = cdd[(long long)k * g.NV + d];
                const double a = avg[(long long)k * g.NV + d];
                if (cn > 0 && a > 0.0) {
                    int gr;
                    double v;
                    entry(g, A, c, k, spos[d], sg[d], gr, v);
                    sum_rex += v * v;
                    sum1 += v * v * v * (10.0 - 15.0 * v + 6.0 * v * v);
                    h_phi_def += (v * v * v * (10.0 - 15.0 * v + 6.0 * v * v)) * (a / cn);
                }
            }
            int count = 0;
            for (int it = 0; it < nG; it++)
                if (gidx[(long long)it * g.NV + d] > P.init_gid) count++;
            for (int it = 0; it < nG; it++) {
                const int index = gidx[(long long)it * g.NV + d];
                if (index < 1 || index > P.n_rho) {
                    atomicOr(err, PFE_GRAIN);
                    break;
                }
                const double s = lookup_t(T, g, A, q0, c, index);
                const double h_phi = s * s * s * (10.0 - 15.0 * s + 6.0 * s * s);
                const double dh_phi = s * s * (30.0 - 60.0 * s + 30.0 * s * s);
                const double center_rex = s;
                // laplacian() (src/pf_xrx.c:6-9): faces summed in the call's
                // order, edges in the order its permuted parameter names sum
                // them (LAPQ)
                double f = 0.0, e = 0.0;
#pragma unroll 1
                for (int q = 0; q < 18; q++) {
                    const int dx = LAPQ[q][0], dy = LAPQ[q][1], dz = LAPQ[q][2];
                    const double v =
                        lookup_t(T, g, A, tq(tx, ty, tz, dx, dy, dz), cix(g, x + dx, y + dy, z + dz), index);
                    if (q < 6) f = q == 0 ? v : f + v;
                    else e = q == 6 ? v : e + v;
                }
                const double lap_rex = (0.5 * f + 0.125 * e) - 4.5 * center_rex;
                const double ra = rho_at(g, keys, gfirst, gend, refs, avg, cdd, rho_in, index, d);
                if (ra > 0.0) {
                    const double cc = center_rex;
                    const double value_rex =
                        cc - P.M_coeff * (P.m * (cc * cc * cc - cc + 3.0 * cc * (sum_rex - cc * cc)) -
                                          P.kappa * lap_rex +
                                          P.E_disl * dh_phi * ((sum1 - h_phi) * ra - (h_phi_def - h_phi * ra)) /
                                              (sum1 * sum1));
                    if (count > 0 && index > P.init_gid && cc < 0.5 && value_rex > 0.5) trig = 1;
                    if (value_rex > 1E-04) ok = ok && t.put(index, value_rex);
                }
            }
        } else if (NN == 1) {
            for (int it = 0; it < nG; it++) {
                const int index = gidx[(long long)it * g.NV + d];
                ok = ok && t.put(index, lookup_t(T, g, A, q0, c, index));
            }
        }
        if (!ok) atomicOr(err, PFE_CAP);
        t.finish();
        trig_out[d] = trig;
    }
}

__global__ __launch_bounds__(NT) void k_diff(long long n, const int *a, const int *b, int *changed)
{
    for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += (long long)gridDim.x * NT)
        if (a[i] != b[i]) *changed = 1;
}

// after the sweep: the resets to 100 and rho_avg of each written grain (its
// last writer)
__global__ __launch_bounds__(NT) void k_finish(Dims g, const int *trig, double *defeng, const unsigned long long *keys, const int *refs,
                         long long nw, const double *avg, const int *cdd, double *rho)
{
    for (long long d = blockIdx.x * (long long)NT + threadIdx.x; d < g.NV; d += (long long)gridDim.x * NT) {
        if (trig[d]) {
            int x, y, z;
            vox(g, d, x, y, z);
            defeng[cix(g, x, y, z)] = 100.0;
        }
    }
    for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < nw; i += (long long)gridDim.x * NT) {
        const unsigned long long gr = keys[i] / (unsigned long long)g.NV;
        if (i + 1 == nw || keys[i + 1] / (unsigned long long)g.NV != gr) rho[gr] = writer_value(g, avg, cdd, refs[i]);
    }
}

__device__ inline int wrap3(int p, int L) { return p < 3 ? p + L : p >= L + 3 ? p - L : p; }

// boundary_conditions on one rank: each halo cell is filled from its periodic
// image entry by entry through the lookup (copies of copies add nothing: a
// built list has values below EPSILON only in its last entry)
__global__ __launch_bounds__(NT) void k_halo(Dims g, Lists B, int *err)
{
    for (long long c = blockIdx.x * (long long)NT + threadIdx.x; c < g.NC; c += (long long)gridDim.x * NT) {
        const int z = (int)(c % g.lz);
        const long long r = c / g.lz;
        const int y = (int)(r % g.ly), x = (int)(r / g.ly);
        if (interior(g, x, y, z)) continue;
        const long long s = cix(g, wrap3(x, g.Lx), wrap3(y, g.Ly), wrap3(z, g.Lz));
        Build t(B, g.NC, c);
        const int n = B.cnt[s];
        bool ok = true;
        for (int j = 0; j < n; j++) ok = ok && t.put(B.idx[(long long)j * g.NC + s], B.val[(long long)j * g.NC + s]);
        if (!ok) atomicOr(err, PFE_CAP);
        t.finish();
    }
}

// mesh_extend on one rank: periodic halos of the stored energy and grains
__global__ __launch_bounds__(NT) void k_mesh(Dims g, double *defeng, int *gid)
{
    for (long long c = blockIdx.x * (long long)NT + threadIdx.x; c < g.NC; c += (long long)gridDim.x * NT) {
        const int z = (int)(c % g.lz);
        const long long r = c / g.lz;
        const int y = (int)(r % g.ly), x = (int)(r / g.ly);
        if (interior(g, x, y, z)) continue;
        const long long s = cix(g, wrap3(x, g.Lx), wrap3(y, g.Ly), wrap3(z, g.Lz));
        defeng[c] = defeng[s];
        gid[c] = gid[s];
    }
}

// Update_gID (src/pf_xrx.c:11-66) and the sum check of GetStatistics
// (src/pf_xrx.c:1364-1386)
__global__ __launch_bounds__(NT) void k_gid(Dims g, Lists A, int *gid_new, int *err)
{
    for (long long d = blockIdx.x * (long long)NT + threadIdx.x; d < g.NV; d += (long long)gridDim.x * NT) {
        int x, y, z;
        vox(g, d, x, y, z);
        const long long c = cix(g, x, y, z);
        const int n = A.cnt[c];
        int id = 0;
        double mx = 0.0, sum = 0.0;
        for (int i = 0; i < n; i++) {
            const double v = A.val[(long long)i * g.NC + c];
            if (mx < v) {
                mx = v;
                id = A.idx[(long long)i * g.NC + c];
            }
            sum += v;
        }
        gid_new[d] = id;
        if (sum < 0.5) atomicOr(err, PFE_SUM);
    }
}

__global__ __launch_bounds__(NT) void k_clear(long long n, int *cnt)
{
    for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += (long long)gridDim.x * NT) cnt[i] = 0;
}

// PF_XRX's initial state (src/pf_xrx.c:1538-1560): one entry (gID, 1.0) per
// voxel, the stored energy and the grain field
__global__ __launch_bounds__(NT) void k_init(Dims g, Lists A, const int *gID, const double *rho, double *defeng, int *gid)
{
    for (long long d = blockIdx.x * (long long)NT + threadIdx.x; d < g.NV; d += (long long)gridDim.x * NT) {
        int x, y, z;
        vox(g, d, x, y, z);
        const long long c = cix(g, x, y, z);
        A.cnt[c] = 1;
        A.idx[c] = gID[d];
        A.val[c] = 1.0;
        defeng[c] = rho[d];
        gid[c] = gID[d];
    }
}

// interior lists from host layout [v*CAP + k]
__global__ __launch_bounds__(NT) void k_load(Dims g, Lists A, const int *n, const int *idx, const double *val, const int *GID,
                       const double *de, double *defeng, int *gid, int *err)
{
    for (long long d = blockIdx.x * (long long)NT + threadIdx.x; d < g.NV; d += (long long)gridDim.x * NT) {
        int x, y, z;
        vox(g, d, x, y, z);
        const long long c = cix(g, x, y, z);
        int m = n[d];
        if (m < 0 || m > CAP) {
            atomicOr(err, PFE_CAP);
            m = 0;
        }
        A.cnt[c] = m;
        for (int k = 0; k < m; k++) {
            A.idx[(long long)k * g.NC + c] = idx[d * CAP + k];
            A.val[(long long)k * g.NC + c] = val[d * CAP + k];
        }
        defeng[c] = de[d];
        gid[c] = GID[d];
    }
}

__global__ __launch_bounds__(NT) void k_store(Dims g, Lists A, const double *defeng, int *n, int *idx, double *val, double *de)
{
    for (long long d = blockIdx.x * (long long)NT + threadIdx.x; d < g.NV; d += (long long)gridDim.x * NT) {
        int x, y, z;
        vox(g, d, x, y, z);
        const long long c = cix(g, x, y, z);
        const int m = A.cnt[c];
        n[d] = m;
        for (int k = 0; k < CAP; k++) {
            idx[d * CAP + k] = k < m ? A.idx[(long long)k * g.NC + c] : 0;
            val[d * CAP + k] = k < m ? A.val[(long long)k * g.NC + c] : 0.0;
        }
        de[d] = defeng[c];
    }
}

int grid_for(long long n)
{
    long long b = (n + NT - 1) / NT;
    if (b > 65536) b = 65536;
    return (int)(b < 1 ? 1 : b);
}

} // namespace

struct evpgpu_pf {
    Dims g;
    int n_rho;
    hipStream_t stream;
    Lists A, B;
    double *defeng;
    int *gid;
    int *gcnt, *gidx, *act, *emask, *nn, *spos, *sg, *wn, *trig0, *trig1, *gid_new;
    long long *woff;
    double *avg;
    int *cdd;
    double *rho_in, *rho_out;
    long long *gfirst, *gend; // each grain's run in the sorted writer keys
    unsigned long long *keys[2];
    int *refs[2];
    long long wcap;
    void *tmp;
    size_t tmp_bytes;
    int *flags; // [0] error bits, [1] changed, [2] voxels on the update branch
    int *h_flags;
    long long *h_total;
    int *h_last;
    hipEvent_t ev0, ev1, eu[3];
    int loaded;
    evpgpu_pf_stats st;
};

namespace {

int pf_err(int bits)
{
    if (bits & PFE_CAP) return EVPGPU_ERR_PF_CAP;
    if (bits & PFE_GRAIN) return EVPGPU_ERR_PF_GRAIN;
    if (bits & PFE_SUM) return EVPGPU_ERR_PF_SUM;
    return EVPGPU_OK;
}

template <typename T> int dalloc(T **p, long long n)
{
    return hipMalloc((void **)p, (size_t)(n > 0 ? n : 1) * sizeof(T)) == hipSuccess ? 0 : 1;
}

int ensure_writers(evpgpu_pf *p, long long nw)
{
    if (nw <= p->wcap) return EVPGPU_OK;
    long long cap = p->wcap * 2;
    while (cap < nw) cap *= 2;
    for (int i = 0; i < 2; i++) {
        (void)hipFree(p->keys[i]);
        (void)hipFree(p->refs[i]);
        p->keys[i] = nullptr;
        p->refs[i] = nullptr;
        if (dalloc(&p->keys[i], cap) || dalloc(&p->refs[i], cap)) return EVPGPU_ERR_NOMEM;
    }
    p->wcap = cap;
    return EVPGPU_OK;
}

int ensure_tmp(evpgpu_pf *p, size_t bytes)
{
    if (bytes <= p->tmp_bytes) return EVPGPU_OK;
    (void)hipFree(p->tmp);
    p->tmp = nullptr;
    if (hipMalloc(&p->tmp, bytes) != hipSuccess) return EVPGPU_ERR_NOMEM;
    p->tmp_bytes = bytes;
    return EVPGPU_OK;
}

int read_flags(evpgpu_pf *p)
{
    EVP_CHECK(hipMemcpyAsync(p->h_flags, p->flags, 3 * sizeof(int), hipMemcpyDeviceToHost, p->stream));
    EVP_CHECK(hipStreamSynchronize(p->stream));
    return EVPGPU_OK;
}

int clear_flag(evpgpu_pf *p, int which)
{
    hipLaunchKernelGGL(k_clear, dim3(1), dim3(NT), 0, p->stream, 1LL, p->flags + which);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

// one substep: Update_grid, boundary_conditions of the new lists, swap,
// Update_gID, mesh_extend, the sum check
int substep(evpgpu_pf *p, const Coef &P)
{
    const Dims &g = p->g;
    hipStream_t s = p->stream;
    const int gv = grid_for(g.NV), gc = grid_for(g.NC);
    int st;
    // tiles of TX x TY x TZ voxels (k_glist, k_avg, k_update)
    const int tiles = ((g.Lx + TX - 1) / TX) * ((g.Ly + TY - 1) / TY) * ((g.Lz + TZ - 1) / TZ);
    hipLaunchKernelGGL(k_glist, dim3(tiles), dim3(NT), 0, s, g, p->A, p->gcnt, p->gidx, p->flags);
    hipLaunchKernelGGL(k_act_init, dim3(gv), dim3(NT), 0, s, g, p->A, p->gcnt, p->gidx, p->act, p->emask);
    const int act_tiles = ((g.Lx + ACT_T - 1) / ACT_T) * ((g.Ly + ACT_T - 1) / ACT_T);
    for (int pass = 0;; pass++) {
        if (pass == MAX_PASSES) return EVPGPU_ERR_HIP;
        if ((st = clear_flag(p, 1))) return st;
        hipLaunchKernelGGL(k_act_iter, dim3(act_tiles), dim3(ACT_NT), 0, s, g, p->emask, p->act, p->flags + 1);
        if ((st = read_flags(p))) return st;
        p->st.activity_passes++;
        if (p->h_flags[0]) return pf_err(p->h_flags[0]);
        if (!p->h_flags[1]) break;
    }
    if ((st = clear_flag(p, 2))) return st;
    hipLaunchKernelGGL(k_slot, dim3(gv), dim3(NT), 0, s, g, p->A, p->gcnt, p->gidx, p->act, p->nn, p->spos, p->sg,
                       p->wn, p->flags + 2, p->flags);
    // writer offsets
    size_t need = 0;
    EVP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, p->wn, p->woff, g.NV, s));
    if ((st = ensure_tmp(p, need))) return st;
    EVP_CHECK(hipcub::DeviceScan::ExclusiveSum(p->tmp, need, p->wn, p->woff, g.NV, s));
    EVP_CHECK(hipMemcpyAsync(p->h_total, p->woff + g.NV - 1, sizeof(long long), hipMemcpyDeviceToHost, s));
    EVP_CHECK(hipMemcpyAsync(p->h_last, p->wn + g.NV - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    EVP_CHECK(hipStreamSynchronize(s));
    const long long nw = p->h_total[0] + p->h_last[0];
    if ((st = ensure_writers(p, nw))) return st;
    hipLaunchKernelGGL(k_emit, dim3(gv), dim3(NT), 0, s, g, p->A, p->nn, p->spos, p->sg, p->wn, p->woff,
                       p->keys[0], p->refs[0], p->n_rho, p->flags);
    int end_bit = 1;
    while (end_bit < 64 && ((unsigned long long)p->n_rho * (unsigned long long)g.NV) >> end_bit) end_bit++;
    if (nw > 0) {
        need = 0;
        EVP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, p->keys[0], p->keys[1], p->refs[0], p->refs[1],
                                                     (int)nw, 0, end_bit, s));
        if ((st = ensure_tmp(p, need))) return st;
        EVP_CHECK(hipcub::DeviceRadixSort::SortPairs(p->tmp, need, p->keys[0], p->keys[1], p->refs[0], p->refs[1],
                                                     (int)nw, 0, end_bit, s));
    }
    p->st.rho_writes += nw;
    hipLaunchKernelGGL(k_clear, dim3(grid_for(2LL * p->n_rho)), dim3(NT), 0, s, 2LL * p->n_rho, (int *)p->gfirst);
    hipLaunchKernelGGL(k_clear, dim3(grid_for(2LL * p->n_rho)), dim3(NT), 0, s, 2LL * p->n_rho, (int *)p->gend);
    if (nw > 0)
        hipLaunchKernelGGL(k_grange, dim3(grid_for(nw)), dim3(NT), 0, s, (const unsigned long long *)p->keys[1], nw,
                           g.NV, p->gfirst, p->gend);
    // trigger fixed point; tout holds the resets the previous pass used
    int *tin = p->trig0, *tout = p->trig1;
    hipLaunchKernelGGL(k_clear, dim3(gv), dim3(NT), 0, s, g.NV, tin);
    for (int pass = 0;; pass++) {
        if (pass == MAX_PASSES) return EVPGPU_ERR_HIP;
        EVP_CHECK(hipEventRecord(p->eu[0], s));
        hipLaunchKernelGGL(k_avg, dim3(tiles), dim3(NT), 0, s, g, p->A, p->wn, p->spos, p->sg, p->gid, p->defeng,
                           tin, pass ? tout : (const int *)nullptr, p->avg, p->cdd);
        EVP_CHECK(hipEventRecord(p->eu[1], s));
        hipLaunchKernelGGL(k_update, dim3(tiles), dim3(NT), 0, s, g, p->A, p->B, P, p->gcnt, p->gidx, p->nn, p->spos,
                           p->sg, p->avg, p->cdd, p->keys[1], p->gfirst, p->gend, p->refs[1], p->rho_in, tout,
                           p->flags);
        EVP_CHECK(hipEventRecord(p->eu[2], s));
        if ((st = clear_flag(p, 1))) return st;
        hipLaunchKernelGGL(k_diff, dim3(gv), dim3(NT), 0, s, g.NV, tin, tout, p->flags + 1);
        if ((st = read_flags(p))) return st;
        p->st.trigger_passes++;
        {
            float a = 0.f, u = 0.f;
            (void)hipEventElapsedTime(&a, p->eu[0], p->eu[1]);
            (void)hipEventElapsedTime(&u, p->eu[1], p->eu[2]);
            p->st.ms_avg += a;
            p->st.ms_update += u;
            p->st.update_launches++;
        }
        if (p->h_flags[0]) return pf_err(p->h_flags[0]);
        int *t = tin;
        tin = tout;
        tout = t;
        if (!p->h_flags[1]) break;
    }
    hipLaunchKernelGGL(k_finish, dim3(gv), dim3(NT), 0, s, g, tin, p->defeng, p->keys[1], p->refs[1], nw, p->avg,
                       p->cdd, p->rho_in);
    hipLaunchKernelGGL(k_halo, dim3(gc), dim3(NT), 0, s, g, p->B, p->flags);
    Lists t = p->A;
    p->A = p->B;
    p->B = t;
    hipLaunchKernelGGL(k_clear, dim3(gc), dim3(NT), 0, s, g.NC, p->B.cnt);
    hipLaunchKernelGGL(k_gid, dim3(gv), dim3(NT), 0, s, g, p->A, p->gid_new, p->flags);
    hipLaunchKernelGGL(k_mesh, dim3(gc), dim3(NT), 0, s, g, p->defeng, p->gid);
    if ((st = read_flags(p))) return st;
    if (p->h_flags[0]) return pf_err(p->h_flags[0]);
    p->st.updated += p->h_flags[2];
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

Coef coef_of(const evpgpu_pf_params *P, int n_rho)
{
    // src/pf_xrx.c:1497-1504
    const double dx = P->pf_length_scale;
    const double L_gb = 3.0 * dx;
    const double dt = P->pf_time_step;
    Coef c;
    c.M_coeff = dt * 4.0 / 3.0 * P->M_gb / L_gb;
    c.kappa = 3.0 / 4.0 * P->E_gb * L_gb / dx / dx;
    c.m = 6 * P->E_gb / L_gb;
    c.E_disl = P->alpha * P->mu * P->bb * P->bb;
    c.init_gid = P->init_gid;
    c.n_rho = n_rho;
    return c;
}

} // namespace

extern "C" int evpgpu_pf_create(evpgpu_pf **out, int Lx, int Ly, int Lz, int n_rho_avg)
{
    if (!out) return EVPGPU_ERR_ARG;
    *out = nullptr;
    if (Lx < 3 || Ly < 3 || Lz < 3 || Lx > 1024 || Ly > 1024 || Lz > 1024 || n_rho_avg < 1)
        return EVPGPU_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return EVPGPU_ERR_NODEV;
    evpgpu_pf *p = (evpgpu_pf *)calloc(1, sizeof(evpgpu_pf));
    if (!p) return EVPGPU_ERR_NOMEM;
    Dims &g = p->g;
    g.Lx = Lx;
    g.Ly = Ly;
    g.Lz = Lz;
    g.lx = Lx + 6;
    g.ly = Ly + 6;
    g.lz = Lz + 6;
    g.NC = (long long)g.lx * g.ly * g.lz;
    g.NV = (long long)Lx * Ly * Lz;
    if ((unsigned long long)g.NV * CAP > 0x7fffffffULL) {
        free(p);
        return EVPGPU_ERR_ARG;
    }
    p->n_rho = n_rho_avg;
    int bad = 0;
    for (Lists *L : {&p->A, &p->B}) {
        bad |= dalloc(&L->cnt, g.NC);
        bad |= dalloc(&L->idx, g.NC * CAP);
        bad |= dalloc(&L->val, g.NC * CAP);
    }
    bad |= dalloc(&p->defeng, g.NC) | dalloc(&p->gid, g.NC);
    bad |= dalloc(&p->gcnt, g.NV) | dalloc(&p->gidx, g.NV * CAP) | dalloc(&p->act, g.NV) | dalloc(&p->nn, g.NV);
    bad |= dalloc(&p->emask, g.NV);
    bad |= dalloc(&p->spos, g.NV) | dalloc(&p->sg, g.NV) | dalloc(&p->wn, g.NV) | dalloc(&p->woff, g.NV);
    bad |= dalloc(&p->trig0, g.NV) | dalloc(&p->trig1, g.NV) | dalloc(&p->gid_new, g.NV);
    bad |= dalloc(&p->avg, g.NV * CAP) | dalloc(&p->cdd, g.NV * CAP);
    bad |= dalloc(&p->rho_in, n_rho_avg) | dalloc(&p->flags, 3);
    bad |= dalloc(&p->gfirst, n_rho_avg) | dalloc(&p->gend, n_rho_avg);
    bad |= hipHostMalloc((void **)&p->h_flags, 3 * sizeof(int), 0) != hipSuccess;
    bad |= hipHostMalloc((void **)&p->h_total, sizeof(long long), 0) != hipSuccess;
    bad |= hipHostMalloc((void **)&p->h_last, sizeof(int), 0) != hipSuccess;
    p->wcap = 0;
    if (!bad) {
        p->wcap = g.NV;
        for (int i = 0; i < 2; i++) bad |= dalloc(&p->keys[i], p->wcap) | dalloc(&p->refs[i], p->wcap);
    }
    bad |= hipStreamCreate(&p->stream) != hipSuccess;
    bad |= hipEventCreate(&p->ev0) != hipSuccess;
    bad |= hipEventCreate(&p->ev1) != hipSuccess;
    for (int i = 0; i < 3; i++) bad |= hipEventCreate(&p->eu[i]) != hipSuccess;
    if (!bad) bad |= hipMemset(p->flags, 0, 3 * sizeof(int)) != hipSuccess;
    if (!bad) bad |= hipMemset(p->A.cnt, 0, g.NC * sizeof(int)) != hipSuccess;
    if (!bad) bad |= hipMemset(p->B.cnt, 0, g.NC * sizeof(int)) != hipSuccess;
    if (bad) {
        evpgpu_pf_destroy(p);
        return EVPGPU_ERR_NOMEM;
    }
    *out = p;
    return EVPGPU_OK;
}

extern "C" int evpgpu_pf_destroy(evpgpu_pf *p)
{
    if (!p) return EVPGPU_ERR_ARG;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    for (Lists *L : {&p->A, &p->B}) {
        (void)hipFree(L->cnt);
        (void)hipFree(L->idx);
        (void)hipFree(L->val);
    }
    void *dev[] = {p->defeng, p->gid, p->gcnt, p->gidx, p->act, p->emask, p->nn,    p->spos,    p->sg,      p->wn,
                   p->woff,   p->trig0, p->trig1, p->gid_new, p->avg, p->cdd, p->rho_in, p->flags, p->keys[0],
                   p->keys[1], p->refs[0], p->refs[1], p->tmp, p->gfirst, p->gend};
    for (void *q : dev) (void)hipFree(q);
    if (p->h_flags) (void)hipHostFree(p->h_flags);
    if (p->h_total) (void)hipHostFree(p->h_total);
    if (p->h_last) (void)hipHostFree(p->h_last);
    if (p->ev0) (void)hipEventDestroy(p->ev0);
    if (p->ev1) (void)hipEventDestroy(p->ev1);
    for (int i = 0; i < 3; i++)
        if (p->eu[i]) (void)hipEventDestroy(p->eu[i]);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    free(p);
    return EVPGPU_OK;
}

extern "C" int evpgpu_pf_set_lists(evpgpu_pf *p, const int *n, const int *idx, const double *val, const int *GID,
                                   const double *defeng)
{
    if (!p || !n || !idx || !val || !GID || !defeng) return EVPGPU_ERR_ARG;
    const Dims &g = p->g;
    int *dn = nullptr, *di = nullptr, *dg = nullptr;
    double *dv = nullptr, *dd = nullptr;
    int st = EVPGPU_OK;
    if (dalloc(&dn, g.NV) || dalloc(&di, g.NV * CAP) || dalloc(&dg, g.NV) || dalloc(&dv, g.NV * CAP) ||
        dalloc(&dd, g.NV))
        st = EVPGPU_ERR_NOMEM;
    hipStream_t s = p->stream;
    if (!st) {
        bool ok = hipMemcpyAsync(dn, n, g.NV * sizeof(int), hipMemcpyHostToDevice, s) == hipSuccess;
        ok = ok && hipMemcpyAsync(di, idx, g.NV * CAP * sizeof(int), hipMemcpyHostToDevice, s) == hipSuccess;
        ok = ok && hipMemcpyAsync(dv, val, g.NV * CAP * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess;
        ok = ok && hipMemcpyAsync(dg, GID, g.NV * sizeof(int), hipMemcpyHostToDevice, s) == hipSuccess;
        ok = ok && hipMemcpyAsync(dd, defeng, g.NV * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess;
        if (!ok) st = EVPGPU_ERR_HIP;
    }
    if (!st) {
        const int gv = grid_for(g.NV), gc = grid_for(g.NC);
        hipLaunchKernelGGL(k_clear, dim3(gc), dim3(NT), 0, s, g.NC, p->A.cnt);
        hipLaunchKernelGGL(k_clear, dim3(gc), dim3(NT), 0, s, g.NC, p->B.cnt);
        hipLaunchKernelGGL(k_load, dim3(gv), dim3(NT), 0, s, g, p->A, dn, di, dv, dg, dd, p->defeng, p->gid,
                           p->flags);
        hipLaunchKernelGGL(k_halo, dim3(gc), dim3(NT), 0, s, g, p->A, p->flags);
        hipLaunchKernelGGL(k_mesh, dim3(gc), dim3(NT), 0, s, g, p->defeng, p->gid);
        st = read_flags(p);
        if (!st && p->h_flags[0]) st = pf_err(p->h_flags[0]);
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(dn);
    (void)hipFree(di);
    (void)hipFree(dg);
    (void)hipFree(dv);
    (void)hipFree(dd);
    p->loaded = st == EVPGPU_OK;
    return st;
}

extern "C" int evpgpu_pf_steps(evpgpu_pf *p, const evpgpu_pf_params *P, int steps, int *gID_new, double *rho_avg)
{
    if (!p || !P || steps < 0 || !rho_avg) return EVPGPU_ERR_ARG;
    if (!p->loaded) return EVPGPU_ERR_STATE;
    if (!(P->pf_length_scale > 0.0)) return EVPGPU_ERR_ARG;
    hipStream_t s = p->stream;
    EVP_CHECK(hipMemcpyAsync(p->rho_in, rho_avg, p->n_rho * sizeof(double), hipMemcpyHostToDevice, s));
    const Coef C = coef_of(P, p->n_rho);
    memset(&p->st, 0, sizeof(p->st));
    hipLaunchKernelGGL(k_clear, dim3(1), dim3(NT), 0, s, 3LL, p->flags);
    EVP_CHECK(hipEventRecord(p->ev0, s));
    int st = EVPGPU_OK;
    for (int it = 0; it < steps && !st; it++) {
        st = substep(p, C);
        p->st.substeps++;
    }
    EVP_CHECK(hipEventRecord(p->ev1, s));
    EVP_CHECK(hipStreamSynchronize(s));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, p->ev0, p->ev1);
    p->st.ms = ms;
    if (st) return st;
    EVP_CHECK(hipMemcpyAsync(rho_avg, p->rho_in, p->n_rho * sizeof(double), hipMemcpyDeviceToHost, s));
    if (gID_new && steps > 0)
        EVP_CHECK(hipMemcpyAsync(gID_new, p->gid_new, p->g.NV * sizeof(int), hipMemcpyDeviceToHost, s));
    EVP_CHECK(hipStreamSynchronize(s));
    return EVPGPU_OK;
}

extern "C" int evpgpu_pf_xrx(evpgpu_pf *p, const evpgpu_pf_params *P, const int *gID, const double *rho, int steps,
                             int *gID_new, double *rho_avg)
{
    if (!p || !P || !gID || !rho || !gID_new || !rho_avg || steps < 0) return EVPGPU_ERR_ARG;
    const Dims &g = p->g;
    hipStream_t s = p->stream;
    int *dg = nullptr;
    double *dr = nullptr;
    if (dalloc(&dg, g.NV) || dalloc(&dr, g.NV)) {
        (void)hipFree(dg);
        return EVPGPU_ERR_NOMEM;
    }
    int st = EVPGPU_OK;
    if (hipMemcpyAsync(dg, gID, g.NV * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dr, rho, g.NV * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess)
        st = EVPGPU_ERR_HIP;
    if (!st) {
        const int gv = grid_for(g.NV), gc = grid_for(g.NC);
        hipLaunchKernelGGL(k_clear, dim3(gc), dim3(NT), 0, s, g.NC, p->A.cnt);
        hipLaunchKernelGGL(k_clear, dim3(gc), dim3(NT), 0, s, g.NC, p->B.cnt);
        hipLaunchKernelGGL(k_init, dim3(gv), dim3(NT), 0, s, g, p->A, dg, dr, p->defeng, p->gid);
        hipLaunchKernelGGL(k_halo, dim3(gc), dim3(NT), 0, s, g, p->A, p->flags);
        hipLaunchKernelGGL(k_mesh, dim3(gc), dim3(NT), 0, s, g, p->defeng, p->gid);
        if (hipGetLastError() != hipSuccess) st = EVPGPU_ERR_HIP;
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(dg);
    (void)hipFree(dr);
    if (st) return st;
    p->loaded = 1;
    // gID_new starts as gID (src/pf_xrx.c:1551) and is rewritten by each substep
    memcpy(gID_new, gID, g.NV * sizeof(int));
    return evpgpu_pf_steps(p, P, steps, gID_new, rho_avg);
}

extern "C" int evpgpu_pf_get_lists(evpgpu_pf *p, int *n, int *idx, double *val, double *defeng)
{
    if (!p) return EVPGPU_ERR_ARG;
    if (!p->loaded) return EVPGPU_ERR_STATE;
    const Dims &g = p->g;
    hipStream_t s = p->stream;
    int *dn = nullptr, *di = nullptr;
    double *dv = nullptr, *dd = nullptr;
    if (dalloc(&dn, g.NV) || dalloc(&di, g.NV * CAP) || dalloc(&dv, g.NV * CAP) || dalloc(&dd, g.NV)) {
        (void)hipFree(dn);
        (void)hipFree(di);
        (void)hipFree(dv);
        return EVPGPU_ERR_NOMEM;
    }
    hipLaunchKernelGGL(k_store, dim3(grid_for(g.NV)), dim3(NT), 0, s, g, p->A, p->defeng, dn, di, dv, dd);
    bool ok = hipGetLastError() == hipSuccess;
    if (ok && n) ok = hipMemcpyAsync(n, dn, g.NV * sizeof(int), hipMemcpyDeviceToHost, s) == hipSuccess;
    if (ok && idx) ok = hipMemcpyAsync(idx, di, g.NV * CAP * sizeof(int), hipMemcpyDeviceToHost, s) == hipSuccess;
    if (ok && val) ok = hipMemcpyAsync(val, dv, g.NV * CAP * sizeof(double), hipMemcpyDeviceToHost, s) == hipSuccess;
    if (ok && defeng) ok = hipMemcpyAsync(defeng, dd, g.NV * sizeof(double), hipMemcpyDeviceToHost, s) == hipSuccess;
    (void)hipStreamSynchronize(s);
    (void)hipFree(dn);
    (void)hipFree(di);
    (void)hipFree(dv);
    (void)hipFree(dd);
    return ok ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

extern "C" int evpgpu_pf_get_stats(evpgpu_pf *p, evpgpu_pf_stats *st)
{
    if (!p || !st) return EVPGPU_ERR_ARG;
    *st = p->st;
    return EVPGPU_OK;
}
