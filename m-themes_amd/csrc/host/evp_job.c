/* evp_job.c -- SetupJob / Evolution / DestroyJob of the drop-in on top of
 * the evpgpu device loop: the power law of the configuration files, or
 * (evp_set_law) the dislocation-density law the reference's build compiles
 * in, with its copper constants, densities and dd.out / drx.out.
 *
 * Setup follows src/init.c:727-1205:
 *   k-space            computed on the fly by the Green kernel (not stored)
 *   phases             ElastStiffnessMatrix + PlasticInit (src/init.c:815-862)
 *   microstructure     InitMicroStruct (src/init.c:233-396); C_gr, its
 *                      inverse and the elastic initial stress depend only on
 *                      the voxel's Euler angles, so they are computed once per
 *                      distinct orientation and shared through an index
 *   boundary conds     src/init.c:864-900
 *   CRSS               src/init.c:1135-1163
 *   InitFields         src/init.c:444-515
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "evp_host.h"

#define PI_REF 3.14159265359 /* src/evp.h:77 (Q8) */

/* law of the jobs opened next (evp_set_law) */
static int G_law = EVP_LAW_POWER;
void evp_set_law(int law) { G_law = law == EVP_LAW_DD ? EVP_LAW_DD : EVP_LAW_POWER; }

struct evp_job {
    evp_namelist nl;
    int law;
    evp_dd_keys dd;
    evpgpu_dd_law dd_law;
    evpgpu_dd_phase dd_ph[EVP_NPHMAX];
    char dir[512];
    int precision;
    evpgpu_ctx *ctx;
    evpgpu_phase ph[EVP_NPHMAX];
    evpgpu_loading ld;
    long long N;
    int *grain, *phase, *orient;
    int n_orient;
    double *sgr;          /* n_orient x 36: LU_inv_66(C_gr) */
    double *sig0;         /* n_orient x 6 : initial elastic stress */
    double *tran0;        /* n_orient x 9 : TranMat_xt2sa */
    double udot[9], scauchy[9], d_vm;
    int creep;
    int last_status;
    evpgpu_macro m0;
    double vm_row[7];
};

/* ---------------------------------------------------- orientation table -- */
typedef struct {
    uint64_t key[4];
    int id;
} okey;

static uint64_t dbits(double x)
{
    uint64_t u;
    memcpy(&u, &x, 8);
    return u;
}

static uint64_t mix(uint64_t h, uint64_t x)
{
    h ^= x + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
    return h * 0xBF58476D1CE4E5B9ULL;
}

/* InitMicroStruct (src/init.c:233-396), single slab */
static int read_microstructure(evp_job *J, const double Cijkl[2][3][3][3][3], double c066[6][6],
                               double *wgt_ph1)
{
    const int Ny = J->nl.CellDim[1], Nz = J->nl.CellDim[2];
    const long long N = J->N;
    char buf[80] = {0};
    FILE *fp = evp_open_rel(J->dir, J->nl.initial_ms);
    if (!fp) return EVP_ERR_FILE;
    long long cap = 1024;
    while (cap < 2 * N) cap <<= 1;
    okey *tab = calloc((size_t)cap, sizeof(okey));
    for (long long i = 0; i < cap; i++) tab[i].id = -1;
    int ocap = 256;
    double *eul = malloc(sizeof(double) * 3 * ocap);
    int *oph = malloc(sizeof(int) * ocap);
    int nor = 0, nph1 = 0;
    double (*cgr)[6][6] = NULL;
    int st = 0;
    for (long long v = 0; v < N; v++) {
        double ph = 0, th = 0, om = 0;
        int ii = 0, jj = 0, kk = 0, gr = 0, jp = 0;
        if (!fgets(buf, 80, fp)) buf[0] = 0;
        sscanf(buf, "%lf %lf %lf %d %d %d %d %d", &ph, &th, &om, &ii, &jj, &kk, &gr, &jp);
        long long idx = ((long long)(ii - 1) * Ny + jj - 1) * Nz + kk - 1;
        if (idx != v || jp < 1 || jp > J->nl.N_phases) {
            st = EVP_ERR_MSINDEX;
            break;
        }
        if (jp == 1) nph1++;
        J->grain[v] = gr;
        J->phase[v] = jp;
        J->orient[v] = -1;
        if (J->nl.Type_phases[jp - 1]) continue; /* gas */
        uint64_t k[4] = {(uint64_t)jp, dbits(ph), dbits(th), dbits(om)};
        uint64_t h = mix(mix(mix(mix(0, k[0]), k[1]), k[2]), k[3]);
        long long slot = (long long)(h & (uint64_t)(cap - 1));
        while (tab[slot].id >= 0 && memcmp(tab[slot].key, k, sizeof(k)) != 0) slot = (slot + 1) & (cap - 1);
        if (tab[slot].id < 0) {
            memcpy(tab[slot].key, k, sizeof(k));
            tab[slot].id = nor;
            if (nor == ocap) {
                ocap *= 2;
                eul = realloc(eul, sizeof(double) * 3 * ocap);
                oph = realloc(oph, sizeof(int) * ocap);
            }
            eul[3 * nor] = ph;
            eul[3 * nor + 1] = th;
            eul[3 * nor + 2] = om;
            oph[nor] = jp;
            nor++;
        }
        J->orient[v] = tab[slot].id;
    }
    fclose(fp);
    free(tab);
    if (!st) {
        if (nor == 0) nor = 1; /* an all-gas grid still needs a valid table row */
        J->n_orient = nor;
        J->sgr = calloc((size_t)nor * 36, sizeof(double));
        J->sig0 = calloc((size_t)nor * 6, sizeof(double));
        J->tran0 = calloc((size_t)nor * 9, sizeof(double));
        cgr = calloc((size_t)nor, sizeof(*cgr));
        for (int o = 0; o < nor && !st; o++) {
            int jp = oph[o];
            double ph = eul[3 * o], th = eul[3 * o + 1], om = eul[3 * o + 2];
            ph *= PI_REF / 180.;
            th *= PI_REF / 180.;
            om *= PI_REF / 180.;
            double sa2xt[3][3], c4[3][3][3][3];
            evp_euler_to_mat(ph, th, om, sa2xt);
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) J->tran0[9 * o + a * 3 + b] = sa2xt[b][a];
            evp_ten4_transform_inv(c4, sa2xt, Cijkl[jp - 1]);
            evp_t4_to_c2(c4, cgr[o]);
            double sg[6][6];
            memcpy(sg, cgr[o], sizeof(sg));
            if ((st = evp_lu_inv66(sg))) break;
            memcpy(J->sgr + 36 * o, sg, sizeof(sg));
        }
        /* C066 = <C_gr> over all voxels, gas counted as zero (Q16) */
        double acc[6][6] = {{0}};
        for (long long v = 0; v < N && !st; v++) {
            int o = J->orient[v];
            if (o < 0) continue;
            for (int a = 0; a < 6; a++)
                for (int b = 0; b < 6; b++) acc[a][b] += cgr[o][a][b];
        }
        for (int a = 0; a < 6; a++)
            for (int b = 0; b < 6; b++) c066[a][b] = acc[a][b] * (1. / N);
        *wgt_ph1 = (1.0 * (double)nph1) / N;
    }
    /* initial elastic stress sigma = C_gr : DisGradAvg (src/init.c:470-484) */
    if (!st) {
        double dga[3][3];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) dga[a][b] = J->udot[a * 3 + b] * J->nl.TimeStep;
        for (int o = 0; o < J->n_orient; o++) {
            double cg[3][3][3][3], s[3][3];
            evp_c2_to_t4(cgr[o], cg);
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    double acc = 0.0;
                    for (int p = 0; p < 3; p++)
                        for (int q = 0; q < 3; q++) acc += cg[a][b][p][q] * dga[p][q];
                    if (J->creep == 1) acc += J->scauchy[a * 3 + b];
                    s[a][b] = acc;
                }
            double *o6 = J->sig0 + 6 * o;
            o6[0] = s[0][0]; o6[1] = s[1][1]; o6[2] = s[2][2];
            o6[3] = s[1][2]; o6[4] = s[0][2]; o6[5] = s[0][1];
        }
    }
    free(cgr);
    free(eul);
    free(oph);
    return st;
}

/* boundary conditions (src/init.c:864-900) */
static int boundary_conditions(evp_job *J)
{
    evp_namelist *nl = &J->nl;
    if ((nl->VelGrad_BC_Flag[0] + nl->VelGrad_BC_Flag[1] + nl->VelGrad_BC_Flag[2]) == 2) return EVP_ERR_BC116;
    double v[6], s[6];
    for (int i = 0; i < 6; i++) {
        v[i] = nl->VelGrad_BC[i] * nl->VelGrad_BC_Flag[i];
        s[i] = nl->Stress_BC[i];
    }
    double U[9] = {v[0], v[5], v[4], v[5], v[1], v[3], v[4], v[3], v[2]};
    double S[9] = {s[0], s[5], s[4], s[5], s[1], s[3], s[4], s[3], s[2]};
    memcpy(J->udot, U, sizeof(U));
    memcpy(J->scauchy, S, sizeof(S));
    double us[3][3], d6[6], d5[5], dev[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) us[a][b] = (U[a * 3 + b] + U[b * 3 + a]) / 2.;
    evp_t2_to_v6(us, d6);
    for (int i = 0; i < 5; i++) d5[i] = d6[i];
    evp_v5_to_t2(d5, dev);
    double dvm = 0.0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) dvm += pow(dev[a][b], 2.0);
    J->d_vm = sqrt(2. / 3. * dvm);
    for (int i = 0; i < 6; i++)
        if ((nl->Stress_BC_Flag[i] * nl->VelGrad_BC_Flag[i] != 0) ||
            ((nl->Stress_BC_Flag[i] + nl->VelGrad_BC_Flag[i]) != 1))
            return EVP_ERR_BC114;
    J->creep = 0;
    for (int i = 0; i < 6; i++) J->creep += nl->Stress_BC_Flag[i];
    J->creep /= 6;
    for (int i = 0; i < 6; i++) {
        J->ld.velgrad_bc[i] = v[i];
        J->ld.stress_bc[i] = s[i];
        J->ld.velgrad_bc_flag[i] = nl->VelGrad_BC_Flag[i];
        J->ld.stress_bc_flag[i] = nl->Stress_BC_Flag[i];
    }
    J->ld.d_vm0 = J->d_vm;
    return 0;
}

/* upload the post-setup state (fields, CRSS, orientations, macros) */
static int upload_initial(evp_job *J)
{
    const long long N = J->N;
    double *sig = calloc((size_t)(6 + 9 + 6 + 6 + 9 + 24 + 1) * N, sizeof(double));
    double *dg = sig + 6 * N, *eps = dg + 9 * N, *edot = eps + 6 * N, *tran = edot + 6 * N;
    double *crss = tran + 9 * N, *gac = crss + 24 * N;
    for (long long v = 0; v < N; v++) {
        int o = J->orient[v], p = J->phase[v] - 1;
        if (o >= 0) {
            for (int c = 0; c < 6; c++) sig[c * N + v] = J->sig0[6 * o + c];
            for (int c = 0; c < 9; c++) tran[c * N + v] = J->tran0[9 * o + c];
            for (int s = 0; s < J->ph[p].nsys; s++) {
                crss[(2 * s) * N + v] = J->ph[p].tau[s][0];
                crss[(2 * s + 1) * N + v] = J->ph[p].tau[s][1];
            }
        }
    }
    int st = evpgpu_upload_state(J->ctx, sig, dg, eps, edot, tran, crss, gac);
    free(sig);
    if (st) return st;
    /* densities of the dislocation-density law (src/init.c:1073-1093) */
    if (J->law == EVP_LAW_DD && (st = evpgpu_set_dd_law(J->ctx, &J->dd_law, J->dd_ph, J->nl.N_phases)))
        return st;
    return evpgpu_set_macro(J->ctx, &J->m0);
}

evp_job *evp_job_open(const char *input_path, int precision, int *status)
{
    evp_job *J = calloc(1, sizeof(evp_job));
    int st = 0;
    J->precision = precision;
    strncpy(J->dir, input_path, sizeof(J->dir) - 1);
    char *sl = strrchr(J->dir, '/');
    if (sl) *sl = 0;
    else strcpy(J->dir, ".");
    if (evp_parse_namelist(input_path, &J->nl) < 0) {
        st = EVP_ERR_FILE;
        goto fail;
    }
    evp_namelist *nl = &J->nl;
    if (nl->N_phases > EVP_NPHMAX || nl->N_phases < 1) {
        st = EVP_ERR_NPH;
        goto fail;
    }
    J->N = (long long)nl->CellDim[0] * nl->CellDim[1] * nl->CellDim[2];
    if (J->N <= 0) {
        st = EVPGPU_ERR_SIZE;
        goto fail;
    }
    /* phases (src/init.c:815-862).  The dislocation-density build (DD_CU)
     * replaces phase I's elastic constants by copper's and zeroes phase II's
     * (src/init.c:821-824, 840-843).                                       */
    J->law = G_law;
    double shear_G = 0.0;
    if (J->law == EVP_LAW_DD) {
        if (evp_parse_dd_keys(input_path, &J->dd) < 0) {
            st = EVP_ERR_FILE;
            goto fail;
        }
        if (!nl->Type_phases[0]) {
            nl->ElastConst_PhaseI[0] = 241.1674 * 1000;
            nl->ElastConst_PhaseI[1] = 89.1989 * 1000;
            nl->ElastConst_PhaseI[2] = 151.9685 * 1000;
        }
        if (nl->N_phases == 2 && !nl->Type_phases[1])
            nl->ElastConst_PhaseII[0] = nl->ElastConst_PhaseII[1] = nl->ElastConst_PhaseII[2] = 0.0;
        shear_G = evp_dd_shear_modulus(nl->ElastConst_PhaseI);
    }
    double Cijkl[2][3][3][3][3];
    memset(Cijkl, 0, sizeof(Cijkl));
    for (int p = 0; p < nl->N_phases; p++) {
        J->ph[p].is_gas = nl->Type_phases[p] != 0;
        if (!J->ph[p].is_gas) {
            evp_cubic_stiffness(p == 0 ? nl->ElastConst_PhaseI : nl->ElastConst_PhaseII, Cijkl[p]);
            if ((st = evp_parse_sx(J->dir, p == 0 ? nl->Slip_PhaseI : nl->Slip_PhaseII, &J->ph[p]))) goto fail;
        }
    }
    if ((st = boundary_conditions(J))) goto fail;
    J->grain = malloc(sizeof(int) * J->N);
    J->phase = malloc(sizeof(int) * J->N);
    J->orient = malloc(sizeof(int) * J->N);
    double c066[6][6], s066[6][6], wgt_ph1 = 0.0;
    if ((st = read_microstructure(J, (const double(*)[3][3][3][3])Cijkl, c066, &wgt_ph1))) goto fail;
    memcpy(s066, c066, sizeof(s066));
    if ((st = evp_lu_inv66(s066))) goto fail;
    double c0[3][3][3][3], s0[3][3][3][3];
    evp_c2_to_t4(c066, c0);
    evp_c2_to_t4(s066, s0);
    evpgpu_loading *ld = &J->ld;
    ld->dt = nl->TimeStep;
    ld->err_tol = nl->Err;
    ld->iter_max = nl->IterMax;
    ld->update_flag = nl->Update_Flag;
    ld->hard_flag = nl->Hard_Flag;
    ld->elast_bc = nl->ElastBC;
    memcpy(ld->c066, c066, sizeof(c066));
    memcpy(ld->s066, s066, sizeof(s066));
    memcpy(ld->c0, c0, sizeof(c0));
    memcpy(ld->s0, s0, sizeof(s0));
    ld->wgt_ph1 = wgt_ph1;
    /* device */
    if ((st = evpgpu_create(&J->ctx, nl->CellDim[0], nl->CellDim[1], nl->CellDim[2], precision))) goto fail;
    if ((st = evpgpu_set_phases(J->ctx, J->ph, nl->N_phases))) goto fail;
    if (J->law == EVP_LAW_DD) {
        for (int p = 0; p < nl->N_phases; p++)
            if (!J->ph[p].is_gas)
                evp_dd_phase_setup(&J->dd, p, &J->ph[p], shear_G, nl->ElastConst_PhaseI[3], &J->dd_ph[p]);
        J->dd_law.T_K = J->dd.T__K;
        J->dd_law.L0 = J->dd.L0;
        J->dd_law.D_0 = J->dd.D_0;
        J->dd_law.rho_ssd_initial = J->dd.rho_SSD_initial;
        J->dd_law.tau_drx = J->dd.tau_DRX;
        J->dd_law.delta_tau_drx = J->dd.DeltaTau_DRX;
    }
    if ((st = evpgpu_set_loading(J->ctx, ld))) goto fail;
    if ((st = evpgpu_upload_microstructure(J->ctx, J->grain, J->phase, J->orient, J->n_orient, J->sgr))) goto fail;
    /* macro state after InitFields */
    memset(&J->m0, 0, sizeof(J->m0));
    J->m0.d_vm = J->d_vm;
    J->m0.e_vm = J->d_vm * nl->TimeStep;
    for (int i = 0; i < 9; i++) J->m0.disgrad_avg[i] = J->udot[i] * nl->TimeStep;
    if ((st = upload_initial(J))) goto fail;
    if (status) *status = 0;
    return J;
fail:
    if (status) *status = st;
    evp_job_close(J);
    return NULL;
}

void evp_job_close(evp_job *J)
{
    if (!J) return;
    if (J->ctx) evpgpu_destroy(J->ctx);
    free(J->grain);
    free(J->phase);
    free(J->orient);
    free(J->sgr);
    free(J->sig0);
    free(J->tran0);
    free(J);
}

evpgpu_ctx *evp_job_ctx(evp_job *J) { return J ? J->ctx : NULL; }

int evp_job_dims(const evp_job *J, int d[3])
{
    if (!J) return EVPGPU_ERR_ARG;
    memcpy(d, J->nl.CellDim, sizeof(J->nl.CellDim));
    return 0;
}

int evp_job_nsteps(const evp_job *J) { return J ? J->nl.N_steps : 0; }

int evp_job_reset(evp_job *J)
{
    if (!J) return EVPGPU_ERR_ARG;
    return upload_initial(J);
}

int evp_job_law(const evp_job *J) { return J ? J->law : -1; }

int evp_job_set_params(evp_job *J, int phase, const double q[5])
{
    if (!J || !q || phase < 0 || phase >= J->nl.N_phases || J->ph[phase].is_gas) return EVPGPU_ERR_ARG;
    evpgpu_phase *ph = &J->ph[phase];
    for (int s = 0; s < ph->nsys; s++) {
        ph->tau[s][0] = q[EVPGPU_P_TAU0];
        ph->tau[s][1] = q[EVPGPU_P_TAU0];
        ph->tau[s][2] = q[EVPGPU_P_TAU1];
        ph->thet[s][0] = q[EVPGPU_P_THETA0];
        ph->thet[s][1] = q[EVPGPU_P_THETA1];
        ph->nrs[s] = q[EVPGPU_P_NRS];
    }
    int st = evpgpu_set_phases(J->ctx, J->ph, J->nl.N_phases);
    if (st) return st;
    return upload_initial(J);
}

int evp_job_get_params(const evp_job *J, int phase, double q[5])
{
    if (!J || !q || phase < 0 || phase >= J->nl.N_phases || J->ph[phase].is_gas) return EVPGPU_ERR_ARG;
    const evpgpu_phase *ph = &J->ph[phase];
    q[EVPGPU_P_TAU0] = ph->tau[0][0];
    q[EVPGPU_P_TAU1] = ph->tau[0][2];
    q[EVPGPU_P_THETA0] = ph->thet[0][0];
    q[EVPGPU_P_THETA1] = ph->thet[0][1];
    q[EVPGPU_P_NRS] = ph->nrs[0];
    return 0;
}

int evp_job_write_field(evp_job *J, int which, const char *prefix, int step)
{
    if (!J || !prefix || which < EVP_FIELD_SIG || which > EVP_FIELD_DISGRAD) return EVPGPU_ERR_ARG;
    const long long N = J->N;
    double *sig = malloc(sizeof(double) * (6 + 9 + 6 + 6 + 6) * N);
    if (!sig) return EVPGPU_ERR_NOMEM;
    double *dg = sig + 6 * N, *eps = dg + 9 * N, *edot = eps + 6 * N, *out = edot + 6 * N;
    int st = evpgpu_download_state(J->ctx, sig, dg, eps, edot, NULL, NULL, NULL, NULL);
    if (!st) {
        /* DisGrad components as the writers pick them: [i][i], [1][2], [0][2], [0][1] */
        const int ij[6] = {0, 4, 8, 5, 2, 1}, ji[6] = {0, 4, 8, 7, 6, 3};
        for (int c = 0; c < 6; c++)
            for (long long v = 0; v < N; v++) {
                double x;
                switch (which) {
                case EVP_FIELD_SIG: x = sig[c * N + v]; break;
                case EVP_FIELD_EPS: x = eps[c * N + v]; break;
                case EVP_FIELD_EDOT: x = edot[c * N + v]; break;
                case EVP_FIELD_ELS:
                    x = (dg[ij[c] * N + v] + dg[ji[c] * N + v]) / 2.0 - eps[c * N + v];
                    if (c < 3) x = dg[ij[c] * N + v] - eps[c * N + v];
                    break;
                default: x = dg[ij[c] * N + v]; break;
                }
                out[c * N + v] = x;
            }
        st = evp_write_iout6(prefix, step, N, out);
    }
    free(sig);
    return st;
}

int evp_job_write_texture(evp_job *J, const char *prefix, int step)
{
    if (!J || !prefix) return EVPGPU_ERR_ARG;
    double *tran = malloc(sizeof(double) * 9 * J->N);
    if (!tran) return EVPGPU_ERR_NOMEM;
    int st = evpgpu_download_state(J->ctx, NULL, NULL, NULL, NULL, tran, NULL, NULL, NULL);
    if (!st) st = evp_write_texture(prefix, step, J->nl.CellDim, tran, J->grain, J->phase);
    free(tran);
    return st;
}

/* The dislocation-density build checks, at the top of every increment, the
 * von Mises strain the previous increment left: above 4 it runs the DRX
 * nucleation check and, from the first nucleus on, the phase-field relaxation
 * (src/evolution.c:850-1316).  Below it that block changes no field the loop
 * reads.  The DRX event itself is not built here: an increment that would
 * start with it is refused instead of run without it. */
static int drx_gate_open(const evp_job *J)
{
    if (J->law != EVP_LAW_DD) return 0;
    evpgpu_macro m;
    evpgpu_get_macro(J->ctx, &m);
    return m.e_vm > 4.0;
}

int evp_job_run(evp_job *J, int n, int forced, int *iters, evpgpu_stats *stats)
{
    if (!J) return EVPGPU_ERR_ARG;
    for (int k = 0; k < n; k++) {
        if (drx_gate_open(J)) {
            J->last_status = EVP_ERR_DRX;
            return EVP_ERR_DRX;
        }
        evpgpu_stats S;
        int st = evpgpu_increment(J->ctx, forced, &S);
        if (st) {
            J->last_status = st;
            return st;
        }
        if (iters) iters[k] = S.iters;
        if (stats) stats[k] = S;
        evpgpu_macro m;
        evpgpu_get_macro(J->ctx, &m);
        double row[7] = {m.time_tot, m.e_vm, m.evmp, m.d_vm, m.dvmp, m.s_vm, m.s_vm1};
        memcpy(J->vm_row, row, sizeof(row));
        if (S.stopped) break;
    }
    return 0;
}

int evp_job_macro(evp_job *J, evpgpu_macro *m) { return J ? evpgpu_get_macro(J->ctx, m) : EVPGPU_ERR_ARG; }

int evp_job_ids(const evp_job *J, int *grain, int *phase)
{
    if (!J) return EVPGPU_ERR_ARG;
    if (grain) memcpy(grain, J->grain, sizeof(int) * J->N);
    if (phase) memcpy(phase, J->phase, sizeof(int) * J->N);
    return 0;
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

int evp_job_grain_list(const evp_job *J, int *ids)
{
    if (!J) return EVPGPU_ERR_ARG;
    int *tmp = malloc(sizeof(int) * (J->N + 1));
    long long n = 0;
    for (long long v = 0; v < J->N; v++)
        if (!J->nl.Type_phases[J->phase[v] - 1]) tmp[n++] = J->grain[v];
    qsort(tmp, (size_t)n, sizeof(int), cmp_int);
    int m = 0;
    for (long long i = 0; i < n; i++)
        if (i == 0 || tmp[i] != tmp[i - 1]) {
            if (ids) ids[m] = tmp[i];
            m++;
        }
    free(tmp);
    return m;
}

int evp_job_setup_info(const evp_job *J, double *c066, double *s066, double *c0, double *s0)
{
    if (!J) return EVPGPU_ERR_ARG;
    if (c066) memcpy(c066, J->ld.c066, sizeof(J->ld.c066));
    if (s066) memcpy(s066, J->ld.s066, sizeof(J->ld.s066));
    if (c0) memcpy(c0, J->ld.c0, sizeof(J->ld.c0));
    if (s0) memcpy(s0, J->ld.s0, sizeof(J->ld.s0));
    return 0;
}

int evp_job_vm_row(const evp_job *J, double row[7])
{
    if (!J) return EVPGPU_ERR_ARG;
    memcpy(row, J->vm_row, sizeof(J->vm_row));
    return 0;
}

/* ------------------------------------------------ the reference's API -- */
static evp_job *G = NULL;
static char G_input[1024];
static int G_abort = 1, G_status = 0, G_precision = EVPGPU_FP64;
static evp_namelist G_nl;
static FILE *fp_vm = NULL, *fp_err = NULL, *fp_dd = NULL, *fp_drx = NULL;

void evp_set_abort_on_error(int on) { G_abort = on; }
int evp_status(void) { return G_status; }
void evp_set_precision(int p) { G_precision = p; }

static void fail_with(int code, const char *msg)
{
    G_status = code;
    fprintf(stderr, "%s (status %d: %s)\n", msg, code, evpgpu_strerror(code));
    if (G_abort) exit(code > 0 ? code : 1);
}

int GetNameList(char **argv)
{
    strncpy(G_input, argv[1], sizeof(G_input) - 1);
    int ok = evp_parse_namelist(G_input, &G_nl);
    if (ok < 0) {
        printf("PE#%d: Input file not found!\n", 0);
        G_status = EVP_ERR_FILE;
        return 1;
    }
    return ok;
}

void PrintNameList(void) { evp_print_namelist(&G_nl); }

void OpenFiles(void)
{
    fp_vm = fopen("vm.out", "w+");
    /* the dislocation-density build's density averages and its DRX log
     * (src/io.c:220-223; no DRX event before e_vm > 4 leaves the header) */
    if (G_law == EVP_LAW_DD) {
        fp_dd = fopen("dd.out", "w+");
        fp_drx = fopen("drx.out", "w+");
    }
    fp_err = fopen("err.out", "w+");
}

void CloseFiles(void)
{
    if (fp_vm) fclose(fp_vm);
    if (fp_err) fclose(fp_err);
    if (fp_dd) fclose(fp_dd);
    if (fp_drx) fclose(fp_drx);
    fp_vm = fp_err = fp_dd = fp_drx = NULL;
}

void SetupJob(void)
{
    int st = 0;
    printf("\n========================================\n");
    printf("------Start setting up the job......------\n");
    G = evp_job_open(G_input, G_precision, &st);
    if (!G) {
        fail_with(st, "SetupJob failed");
        return;
    }
    if (fp_vm) fprintf(fp_vm, "TIME\tEVM\tEVMP\tDVM\tDVMP\tSVM\tSVMSVM_1\n");
    /* src/init.c:498-502: the dd.out header names 7 of its 9 columns */
    if (fp_dd) fprintf(fp_dd, "TIME\tEVM\tDD_F\tDD_P\tDD_s\tDD_m\tDD_g\n");
    if (fp_drx) fprintf(fp_drx, "TIME\tEVM\tNEW_IDs\n");
    if (fp_err) fprintf(fp_err, "IT\tERRE\tERRS\tSVM\n");
    printf("The initial microstructure contains %d grains.\n", evp_job_grain_list(G, NULL));
    printf("\n-----------Job set up-----------------\n");
    printf("========================================\n");
}

void Evolution(void)
{
    if (!G) {
        fail_with(EVPGPU_ERR_STATE, "Evolution called before SetupJob");
        return;
    }
    printf("\n\n======================================\n");
    printf("-------------Simulation starts----------------\n");
    for (int istep = 0; istep < G->nl.N_steps; istep++) {
        printf("\n****************************************\n");
        printf("STEP = %d\n", istep);
        if (G->nl.N_steps != 1 && fp_err) fprintf(fp_err, "STEP = %d\n", istep);
        if (drx_gate_open(G)) {
            fail_with(EVP_ERR_DRX, "DRX nucleation (e_vm > 4, src/evolution.c:850) is not built");
            return;
        }
        evpgpu_stats S;
        int st = evpgpu_increment(G->ctx, 0, &S);
        if (st) {
            fail_with(st, "Evolution failed");
            return;
        }
        if (S.stopped) break;
        evpgpu_macro m;
        evpgpu_get_macro(G->ctx, &m);
        printf("Strain field error = %f\n", m.err_e);
        printf("Stress field error = %e\n", m.err_s);
        printf("Green iterations = %d\n", S.iters);
        if (fp_vm) {
            fprintf(fp_vm, "%e\t%e\t%e\t%e\t%e\t%e\t%e\n", m.time_tot, m.e_vm, m.evmp, m.d_vm, m.dvmp,
                    m.s_vm, m.s_vm1);
            fflush(fp_vm);
        }
        if (fp_dd && G->law == EVP_LAW_DD) {
            /* src/evolution.c:1565-1570: TimeTot, e_vm and AvgDislocation */
            double a[7];
            if ((st = evpgpu_dd_averages(G->ctx, a))) {
                fail_with(st, "dd.out averages failed");
                return;
            }
            fprintf(fp_dd, "%e\t%e\t%e\t%e\t%e\t%e\t%e\t%e\t%e\n", m.time_tot, m.e_vm, a[0], a[1], a[2],
                    a[3], a[4], a[5], a[6]);
            fflush(fp_dd);
        }
        double row[7] = {m.time_tot, m.e_vm, m.evmp, m.d_vm, m.dvmp, m.s_vm, m.s_vm1};
        memcpy(G->vm_row, row, sizeof(row));
        /* field dumps of the live loop (src/evolution.c:1633-1648): the
         * elastic strain and the stress every 500 increments, under the
         * file names the reference uses there */
        if (G->nl.PrintControl[0] == 1 && istep % 500 == 0) {
            if ((st = evp_job_write_field(G, EVP_FIELD_ELS, "els", istep)) ||
                (st = evp_job_write_field(G, EVP_FIELD_SIG, "sig_afterDRX", istep))) {
                fail_with(st, "field dump failed");
                return;
            }
        }
    }
    printf("-------------Simulation ends----------------\n");
    printf("======================================\n\n");
}

void DestroyJob(void)
{
    evp_job_close(G);
    G = NULL;
}
