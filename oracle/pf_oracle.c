/*
 * pf_oracle.c -- CPU restatement of the one-rank sparse phase-field
 * recrystallisation step of M-THEMES (src/pf_xrx.c, src/PFgrid.h).
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP phase-field kernels):
 * nothing in the product links or calls it.
 *
 * Padded grid: lx = Lx + 6 (three halo planes each side), cell (x, y, z) at
 * (x*ly + y)*lz + z, interior x in [3, lx-3); voxel d = ((x-3)*Ly + y-3)*Lz
 * + z-3 of the unpadded field -- the layouts of src/pf_xrx.c:1523-1560.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pf_oracle.h"

#define EPS 5.0E-3 /* src/PFgrid.h:50 */

typedef struct {
    int n;
    int idx[ORC_PF_CAP];
    double val[ORC_PF_CAP];
} plist;

typedef struct {
    int Lx, Ly, Lz, lx, ly, lz;
    plist *gd, *tmp;  /* grid_rex and Update_grid's temp_rex            */
    double *defeng;   /* DefEng_buf (padded)                              */
    int *gid;         /* GID (padded)                                     */
    int err;
    double dummy;
} pfw;

static long cix(const pfw *W, int x, int y, int z) { return ((long)x * W->ly + y) * W->lz + z; }

/* PFdata::operator[] (src/PFgrid.h:69-87): the entry of grain g; an absent
 * grain re-labels the last entry when its value is below EPSILON, else a
 * zero entry is appended */
int orc_pf_lookup(int *n, int cap, int *idx, double *val, int g)
{
    for (int i = 0; i < *n; i++)
        if (idx[i] == g) return i;
    if (*n > 0 && val[*n - 1] < EPS) {
        idx[*n - 1] = g;
        return *n - 1;
    }
    if (*n >= cap) return -1;
    idx[*n] = g;
    val[*n] = 0.0;
    (*n)++;
    return *n - 1;
}

static double *ref(pfw *W, plist *L, int g)
{
    const int k = orc_pf_lookup(&L->n, ORC_PF_CAP, L->idx, L->val, g);
    if (k < 0) {
        W->err = ORC_PF_ERR_CAP;
        W->dummy = 0.0;
        return &W->dummy;
    }
    return &L->val[k];
}

/* dst[index] = value for each entry of src in order: the halo fill of
 * boundary_conditions, entry by entry through the lookup */
static void lcopy(pfw *W, plist *dst, const plist *src)
{
    for (int j = 0; j < src->n; j++) {
        const double value = src->val[j];
        const int index = src->idx[j];
        *ref(W, dst, index) = value;
    }
}

/* laplacian() (src/pf_xrx.c:6-9): a[] holds the call's arguments in order
 * (right, left, front, back, up, down, left_up, left_down, left_front,
 * left_back, center_up_front, center_up_back, center_down_front,
 * center_down_back, right_up, right_down, right_front, right_back, c); the
 * function names its parameters 7-18 differently, so its edge sum runs over
 * arguments 11, 13, 12, 14, 8, 7, 9, 10, 15, 18, 16, 17 (1-based). */
double orc_pf_laplacian(const double a[19])
{
    return (0.5 * (a[0] + a[1] + a[2] + a[3] + a[4] + a[5]) +
            0.125 * (a[10] + a[12] + a[11] + a[13] + a[7] + a[6] + a[8] + a[9] + a[14] + a[17] +
                     a[15] + a[16]) -
            4.5 * a[18]);
}

/* boundary_conditions(gd, 0, 1, ...) (src/pf_xrx.c:253-681) with one rank:
 * y halos of the interior-x, interior-z cells, then z halos of the
 * interior-x cells over every y, then the x halo planes (planes 3..5 to
 * lx-3..lx-1 and lx-6..lx-4 to 0..2 through the rank's own send/receive) */
static void boundary_conditions(pfw *W, plist *g)
{
    const int lx = W->lx, ly = W->ly, lz = W->lz;
    for (int x = 3; x < lx - 3; x++)
        for (int z = 3; z < lz - 3; z++) {
            lcopy(W, &g[cix(W, x, 0, z)], &g[cix(W, x, ly - 6, z)]);
            lcopy(W, &g[cix(W, x, 1, z)], &g[cix(W, x, ly - 5, z)]);
            lcopy(W, &g[cix(W, x, 2, z)], &g[cix(W, x, ly - 4, z)]);
            lcopy(W, &g[cix(W, x, ly - 3, z)], &g[cix(W, x, 3, z)]);
            lcopy(W, &g[cix(W, x, ly - 2, z)], &g[cix(W, x, 4, z)]);
            lcopy(W, &g[cix(W, x, ly - 1, z)], &g[cix(W, x, 5, z)]);
        }
    for (int x = 3; x < lx - 3; x++)
        for (int y = 0; y < ly; y++) {
            lcopy(W, &g[cix(W, x, y, lz - 1)], &g[cix(W, x, y, 5)]);
            lcopy(W, &g[cix(W, x, y, lz - 2)], &g[cix(W, x, y, 4)]);
            lcopy(W, &g[cix(W, x, y, lz - 3)], &g[cix(W, x, y, 3)]);
            lcopy(W, &g[cix(W, x, y, 0)], &g[cix(W, x, y, lz - 6)]);
            lcopy(W, &g[cix(W, x, y, 1)], &g[cix(W, x, y, lz - 5)]);
            lcopy(W, &g[cix(W, x, y, 2)], &g[cix(W, x, y, lz - 4)]);
        }
    for (int p = 0; p < 3; p++)
        for (int y = 0; y < ly; y++)
            for (int z = 0; z < lz; z++) lcopy(W, &g[cix(W, lx - 3 + p, y, z)], &g[cix(W, 3 + p, y, z)]);
    for (int p = 0; p < 3; p++)
        for (int y = 0; y < ly; y++)
            for (int z = 0; z < lz; z++) lcopy(W, &g[cix(W, p, y, z)], &g[cix(W, lx - 6 + p, y, z)]);
}

/* mesh_extend (src/pf_xrx.c:683-893) with one rank: every halo cell of
 * DefEng_buf and GID takes the value of its periodic image */
static int wrap(int p, int L) { return p < 3 ? p + L : p >= L + 3 ? p - L : p; }
static void mesh_extend(pfw *W)
{
    for (int x = 0; x < W->lx; x++)
        for (int y = 0; y < W->ly; y++)
            for (int z = 0; z < W->lz; z++) {
                const int xs = wrap(x, W->Lx), ys = wrap(y, W->Ly), zs = wrap(z, W->Lz);
                if (xs == x && ys == y && zs == z) continue;
                W->defeng[cix(W, x, y, z)] = W->defeng[cix(W, xs, ys, zs)];
                W->gid[cix(W, x, y, z)] = W->gid[cix(W, xs, ys, zs)];
            }
}

void orc_pf_coeffs(const orc_pf_params *P, double out[4])
{
    /* src/pf_xrx.c:1497-1504 */
    const double dx = P->pf_length_scale;
    const double L_gb = 3.0 * dx;
    const double dt = P->pf_time_step;
    out[0] = dt * 4.0 / 3.0 * P->M_gb / L_gb;            /* M_coeff */
    out[1] = 3.0 / 4.0 * P->E_gb * L_gb / dx / dx;       /* kappa   */
    out[2] = 6 * P->E_gb / L_gb;                         /* m       */
    out[3] = P->alpha * P->mu * P->bb * P->bb;           /* E_disl  */
}

/* Update_grid (src/pf_xrx.c:895-1311) */
static void update_grid(pfw *W, const orc_pf_params *P, double *rho_avg, int n_rho_avg)
{
    double C[4];
    orc_pf_coeffs(P, C);
    const double M_coeff = C[0], kappa = C[1], m = C[2], E_disl = C[3];
    const int init_gid = P->init_gid;
    const int lx = W->lx, ly = W->ly, lz = W->lz;
    plist *gd = W->gd, *tmp = W->tmp;
    int G[27 * ORC_PF_CAP];
    for (int x = 3; x < lx - 3 && !W->err; ++x)
        for (int y = 3; y < ly - 3 && !W->err; ++y)
            for (int z = 3; z < lz - 3 && !W->err; ++z) {
                const long pidx_b = cix(W, x, y, z);
                const int xmin = x - 1, xmax = x + 1, ymin = y - 1, ymax = y + 1, zmin = z - 1, zmax = z + 1;
                /* grains above EPSILON anywhere in the 3x3x3 neighbourhood,
                   in order of first appearance */
                int nG = 0;
                for (int i = xmin; i <= xmax; ++i)
                    for (int j = ymin; j <= ymax; ++j)
                        for (int mm = zmin; mm <= zmax; ++mm) {
                            const plist *d = &gd[cix(W, i, j, mm)];
                            for (int k = 0; k < d->n; ++k)
                                if (d->val[k] > EPS) {
                                    const int index = d->idx[k];
                                    int it = 0;
                                    while (it < nG && G[it] != index) it++;
                                    if (it == nG) G[nG++] = index;
                                }
                        }
                plist *upd = &tmp[pidx_b];
                plist *right = &gd[cix(W, xmax, y, z)], *left = &gd[cix(W, xmin, y, z)];
                plist *front = &gd[cix(W, x, ymax, z)], *back = &gd[cix(W, x, ymin, z)];
                plist *up = &gd[cix(W, x, y, zmax)], *down = &gd[cix(W, x, y, zmin)];
                plist *self = &gd[pidx_b];
                plist *left_up = &gd[cix(W, xmin, y, zmax)], *left_down = &gd[cix(W, xmin, y, zmin)];
                plist *left_front = &gd[cix(W, xmin, ymin, z)], *left_back = &gd[cix(W, xmin, ymax, z)];
                plist *cuf = &gd[cix(W, x, ymin, zmax)], *cub = &gd[cix(W, x, ymax, zmax)];
                plist *cdf = &gd[cix(W, x, ymin, zmin)], *cdb = &gd[cix(W, x, ymax, zmin)];
                plist *right_up = &gd[cix(W, xmax, y, zmax)], *right_down = &gd[cix(W, xmax, y, zmin)];
                plist *right_front = &gd[cix(W, xmax, ymin, z)], *right_back = &gd[cix(W, xmax, ymax, z)];
                const int NN = self->n;
                int count = 0;
                if (NN > 1) {
                    double Avg[ORC_PF_CAP];
                    int cdd[ORC_PF_CAP];
                    for (int k = 0; k < NN; k++) {
                        Avg[k] = 0.0;
                        cdd[k] = 0;
                    }
                    /* mean stored energy of each of the voxel's grains over
                       the 7x7x7 box */
                    for (int i = x - 3; i <= x + 3; ++i)
                        for (int j = y - 3; j <= y + 3; ++j)
                            for (int mm = z - 3; mm <= z + 3; ++mm) {
                                const long p = cix(W, i, j, mm);
                                for (int k = 0; k < NN; k++)
                                    if (W->gid[p] == self->idx[k]) {
                                        Avg[k] += W->defeng[p];
                                        cdd[k] += 1;
                                    }
                            }
                    for (int k = 0; k < NN; k++) {
                        const int index = self->idx[k];
                        if (index < 1 || index > n_rho_avg) {
                            W->err = ORC_PF_ERR_GRAIN;
                            return;
                        }
                        rho_avg[index - 1] = 0.0;
                        if (cdd[k] > 0) rho_avg[index - 1] = Avg[k] / cdd[k];
                    }
                    double sum_rex = 0.0, sum1 = 0.0, h_phi_def = 0.0;
                    for (int k = 0; k < NN; k++)
                        if (cdd[k] > 0 && Avg[k] > 0.0) {
                            const double v = self->val[k];
                            sum_rex += v * v;
                            sum1 += v * v * v * (10.0 - 15.0 * v + 6.0 * v * v);
                            h_phi_def += (v * v * v * (10.0 - 15.0 * v + 6.0 * v * v)) * (Avg[k] / cdd[k]);
                        }
                    for (int it = 0; it < nG; it++)
                        if (G[it] > init_gid) count++;
                    for (int it = 0; it < nG; it++) {
                        const int index = G[it];
                        if (index < 1 || index > n_rho_avg) {
                            W->err = ORC_PF_ERR_GRAIN;
                            return;
                        }
                        const double s = *ref(W, self, index);
                        const double h_phi = s * s * s * (10.0 - 15.0 * s + 6.0 * s * s);
                        const double dh_phi = s * s * (30.0 - 60.0 * s + 30.0 * s * s);
                        const double center_rex = *ref(W, self, index);
                        double a[19];
                        a[0] = *ref(W, right, index);
                        a[1] = *ref(W, left, index);
                        a[2] = *ref(W, front, index);
                        a[3] = *ref(W, back, index);
                        a[4] = *ref(W, up, index);
                        a[5] = *ref(W, down, index);
                        a[6] = *ref(W, left_up, index);
                        a[7] = *ref(W, left_down, index);
                        a[8] = *ref(W, left_front, index);
                        a[9] = *ref(W, left_back, index);
                        a[10] = *ref(W, cuf, index);
                        a[11] = *ref(W, cub, index);
                        a[12] = *ref(W, cdf, index);
                        a[13] = *ref(W, cdb, index);
                        a[14] = *ref(W, right_up, index);
                        a[15] = *ref(W, right_down, index);
                        a[16] = *ref(W, right_front, index);
                        a[17] = *ref(W, right_back, index);
                        a[18] = center_rex;
                        const double lap_rex = orc_pf_laplacian(a);
                        const double ra = rho_avg[index - 1];
                        if (ra > 0.0) {
                            const double c = center_rex;
                            const double value_rex =
                                c - M_coeff * (m * (c * c * c - c + 3.0 * c * (sum_rex - c * c)) - kappa * lap_rex +
                                               E_disl * dh_phi * ((sum1 - h_phi) * ra - (h_phi_def - h_phi * ra)) /
                                                   (sum1 * sum1));
                            if (count > 0 && index > init_gid && c < 0.5 && value_rex > 0.5)
                                W->defeng[pidx_b] = 100.0;
                            if (value_rex > 1E-04) *ref(W, upd, index) = value_rex;
                        }
                    }
                }
                if (NN == 1)
                    for (int it = 0; it < nG; it++) {
                        const double v = *ref(W, self, G[it]);
                        *ref(W, upd, G[it]) = v;
                    }
            }
    if (W->err) return;
    boundary_conditions(W, tmp);
    W->gd = tmp;
    W->tmp = gd;
    const long nc = (long)lx * ly * lz;
    for (long i = 0; i < nc; i++) W->tmp[i].n = 0;
}

/* Update_gID (src/pf_xrx.c:11-66): the grain of the first largest value */
static void update_gid(pfw *W, int *gID)
{
    for (int px = 3; px < W->lx - 3; px++)
        for (int py = 3; py < W->ly - 3; py++)
            for (int pz = 3; pz < W->lz - 3; pz++) {
                const long d = ((long)(px - 3) * W->Ly + py - 3) * W->Lz + pz - 3;
                const plist *L = &W->gd[cix(W, px, py, pz)];
                int idx_rex = 0;
                double max_val = 0.0;
                for (int i = 0; i < L->n; i++)
                    if (max_val < L->val[i]) {
                        max_val = L->val[i];
                        idx_rex = L->idx[i];
                    }
                gID[d] = idx_rex;
            }
}

/* the assert of GetStatistics (src/pf_xrx.c:1364-1386) */
static int check_sums(pfw *W)
{
    for (int px = 3; px < W->lx - 3; px++)
        for (int py = 3; py < W->ly - 3; py++)
            for (int pz = 3; pz < W->lz - 3; pz++) {
                const plist *L = &W->gd[cix(W, px, py, pz)];
                double sum = 0.0;
                for (int i = 0; i < L->n; i++) sum += L->val[i];
                if (sum < 0.5) return ORC_PF_ERR_SUM;
            }
    return ORC_PF_OK;
}

static int pfw_open(pfw *W, int Lx, int Ly, int Lz)
{
    memset(W, 0, sizeof(*W));
    if (Lx < 3 || Ly < 3 || Lz < 3) return ORC_PF_ERR_ARG;
    W->Lx = Lx;
    W->Ly = Ly;
    W->Lz = Lz;
    W->lx = Lx + 6;
    W->ly = Ly + 6;
    W->lz = Lz + 6;
    const long nc = (long)W->lx * W->ly * W->lz;
    W->gd = (plist *)calloc(nc, sizeof(plist));
    W->tmp = (plist *)calloc(nc, sizeof(plist));
    W->defeng = (double *)calloc(nc, sizeof(double));
    W->gid = (int *)calloc(nc, sizeof(int));
    if (!W->gd || !W->tmp || !W->defeng || !W->gid) return ORC_PF_ERR_ARG;
    return ORC_PF_OK;
}

static void pfw_close(pfw *W)
{
    free(W->gd);
    free(W->tmp);
    free(W->defeng);
    free(W->gid);
}

static int run_steps(pfw *W, const orc_pf_params *P, int steps, double *rho_avg, int n_rho_avg, int *gID_new,
                     int *n_out, int *idx_out, double *val_out, double *defeng_out)
{
    /* src/pf_xrx.c:1619-1629 */
    for (int it = 0; it < steps; it++) {
        update_grid(W, P, rho_avg, n_rho_avg);
        if (W->err) return W->err;
        update_gid(W, gID_new);
        mesh_extend(W);
        const int st = check_sums(W);
        if (st) return st;
    }
    for (int px = 3; px < W->lx - 3; px++)
        for (int py = 3; py < W->ly - 3; py++)
            for (int pz = 3; pz < W->lz - 3; pz++) {
                const long d = ((long)(px - 3) * W->Ly + py - 3) * W->Lz + pz - 3;
                const long c = cix(W, px, py, pz);
                const plist *L = &W->gd[c];
                if (n_out) n_out[d] = L->n;
                for (int k = 0; k < ORC_PF_CAP; k++) {
                    if (idx_out) idx_out[d * ORC_PF_CAP + k] = k < L->n ? L->idx[k] : 0;
                    if (val_out) val_out[d * ORC_PF_CAP + k] = k < L->n ? L->val[k] : 0.0;
                }
                if (defeng_out) defeng_out[d] = W->defeng[c];
            }
    return ORC_PF_OK;
}

int orc_pf_xrx(int Lx, int Ly, int Lz, const orc_pf_params *P, const int *gID, const double *rho, int steps,
               double *rho_avg, int n_rho_avg, int *gID_new, int *n_out, int *idx_out, double *val_out,
               double *defeng_out)
{
    pfw W;
    int st = pfw_open(&W, Lx, Ly, Lz);
    if (st) {
        pfw_close(&W);
        return st;
    }
    /* src/pf_xrx.c:1538-1560 */
    for (int px = 3; px < W.lx - 3; px++)
        for (int py = 3; py < W.ly - 3; py++)
            for (int pz = 3; pz < W.lz - 3; pz++) {
                const long d = ((long)(px - 3) * Ly + py - 3) * Lz + pz - 3;
                const long b = cix(&W, px, py, pz);
                *ref(&W, &W.gd[b], gID[d]) = 1.0;
                W.defeng[b] = rho[d];
                gID_new[d] = gID[d];
                W.gid[b] = gID[d];
            }
    boundary_conditions(&W, W.gd);
    mesh_extend(&W);
    st = W.err ? W.err : run_steps(&W, P, steps, rho_avg, n_rho_avg, gID_new, n_out, idx_out, val_out, defeng_out);
    pfw_close(&W);
    return st;
}

int orc_pf_run_lists(int Lx, int Ly, int Lz, const orc_pf_params *P, const int *n_in, const int *idx_in,
                     const double *val_in, const int *GID, const double *defeng, int steps, double *rho_avg,
                     int n_rho_avg, int *gID_new, int *n_out, int *idx_out, double *val_out, double *defeng_out)
{
    pfw W;
    int st = pfw_open(&W, Lx, Ly, Lz);
    if (st) {
        pfw_close(&W);
        return st;
    }
    for (int px = 3; px < W.lx - 3 && !st; px++)
        for (int py = 3; py < W.ly - 3 && !st; py++)
            for (int pz = 3; pz < W.lz - 3; pz++) {
                const long d = ((long)(px - 3) * Ly + py - 3) * Lz + pz - 3;
                const long b = cix(&W, px, py, pz);
                plist *L = &W.gd[b];
                if (n_in[d] < 0 || n_in[d] > ORC_PF_CAP) {
                    st = ORC_PF_ERR_ARG;
                    break;
                }
                L->n = n_in[d];
                for (int k = 0; k < L->n; k++) {
                    L->idx[k] = idx_in[d * ORC_PF_CAP + k];
                    L->val[k] = val_in[d * ORC_PF_CAP + k];
                }
                W.defeng[b] = defeng[d];
                W.gid[b] = GID[d];
                gID_new[d] = GID[d];
            }
    if (!st) {
        boundary_conditions(&W, W.gd);
        mesh_extend(&W);
        st = W.err ? W.err
                   : run_steps(&W, P, steps, rho_avg, n_rho_avg, gID_new, n_out, idx_out, val_out, defeng_out);
    }
    pfw_close(&W);
    return st;
}

/* ------------------------------------------------------------ nucleation */

void orc_gb_indicator(int Lx, int Ly, int Lz, const int *grain, int *gb)
{
    /* src/constitutive.c:790-853: the six checks compare with the +x, -x, +y,
       -y, +z neighbours and, for the sixth, -y again (GB_checkZ[1] reads
       pIDX_ym); src/evolution.c:816-828: GB_indicator = 1 when any fails */
    for (int x = 0; x < Lx; x++)
        for (int y = 0; y < Ly; y++)
            for (int z = 0; z < Lz; z++) {
                const long v = ((long)x * Ly + y) * Lz + z;
                const int xp = (x + 1) % Lx, xm = (x + Lx - 1) % Lx, yp = (y + 1) % Ly, ym = (y + Ly - 1) % Ly;
                const int zp = (z + 1) % Lz;
                const int g = grain[v];
                const int chk = (g == grain[((long)xp * Ly + y) * Lz + z]) + (g == grain[((long)xm * Ly + y) * Lz + z]) +
                                (g == grain[((long)x * Ly + yp) * Lz + z]) + (g == grain[((long)x * Ly + ym) * Lz + z]) +
                                (g == grain[((long)x * Ly + y) * Lz + zp]) + (g == grain[((long)x * Ly + ym) * Lz + z]);
                gb[v] = chk < 6 ? 1 : 0;
            }
}

int orc_drx_nucleate(int Lx, int Ly, int Lz, int nsys, const int *grain_in, const double *rho_s,
                     const double *rho_m, const double *g1, const double *g2, const double *g3, const double *kappa,
                     int n_grains, int nucleus_radius, double zeta_kc, double k_c, int *grain_out, int *gid_rex,
                     int *nuclei, double *kc, double *rho_tot, long long *nucleus_vox, int max_new, int *n_new)
{
    const long N = (long)Lx * Ly * Lz;
    int *gb = (int *)malloc(N * sizeof(int));
    if (!gb) return ORC_PF_ERR_ARG;
    orc_gb_indicator(Lx, Ly, Lz, grain_in, gb);
    memcpy(grain_out, grain_in, N * sizeof(int));
    int cnt = 0, st = ORC_PF_OK;
    /* src/integrated.c:181-193 (NUCL_RHO_DIFF off) and 270-300 */
    for (long v = 0; v < N; v++) {
        double r = 0.0;
        for (int i = 0; i < nsys; i++)
            r += rho_s[i * N + v] + rho_m[i * N + v] +
                 sqrt(pow(g1[i * N + v], 2.0) + pow(g2[i * N + v], 2.0) + pow(g3[i * N + v], 2.0));
        rho_tot[v] = r;
        const double diff_rho = r / 12;
        nuclei[v] = 0;
        if (gb[v] > 0 && diff_rho > kappa[v]) {
            gid_rex[v] = 1;
            nuclei[v] = 1;
            kc[v] = zeta_kc * k_c;
            if (cnt < max_new) nucleus_vox[cnt] = v;
            else st = ORC_PF_ERR_ARG;
            cnt++;
        }
    }
    *n_new = cnt;
    if (cnt > 0 && st == ORC_PF_OK) {
        /* src/integrated.c:318-405: consecutive IDs after the existing grains,
           then each nucleus in order paints the voxels within
           |r - r_n|^2 <= nucleus_radius (no periodic wrap) */
        for (int k = 0; k < cnt; k++) grain_out[nucleus_vox[k]] = n_grains + 1 + k;
        int h = 0;
        while ((h + 1) * (h + 1) <= nucleus_radius) h++;
        for (int k = 0; k < cnt; k++) {
            const long v = nucleus_vox[k];
            const int X = (int)(v / ((long)Ly * Lz)), Y = (int)((v / Lz) % Ly), Z = (int)(v % Lz);
            for (int x = X - h; x <= X + h; x++)
                for (int y = Y - h; y <= Y + h; y++)
                    for (int z = Z - h; z <= Z + h; z++) {
                        if (x < 0 || x >= Lx || y < 0 || y >= Ly || z < 0 || z >= Lz) continue;
                        if (pow(x - X, 2) + pow(y - Y, 2) + pow(z - Z, 2) <= nucleus_radius) {
                            const long u = ((long)x * Ly + y) * Lz + z;
                            grain_out[u] = n_grains + 1 + k;
                            gid_rex[u] = 1;
                            nuclei[u] = 1;
                        }
                    }
        }
    }
    free(gb);
    return st;
}
