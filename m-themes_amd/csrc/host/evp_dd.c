/* evp_dd.c -- set-up of the dislocation-density law (the build the reference
 * ships: src/evpFlags.h DD_BASED_FLAG, DD_CU, FORWARD_PLUS_BACKWARD,
 * DD_GND_FIRST_ORDER): its namelist keys (src/io.c:47-93), the per-phase
 * Burgers lengths and sense vectors (src/init.c:179-203), the FCC
 * interaction-strength matrix (src/init.c:593-702) and the projection tables
 * (src/init.c:1108-1125).                                                   */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "evp_host.h"

/* The keys, read with the namelist's rules (200-byte lines, '#' comments,
 * the first definition wins).  Absent keys read as 0.                      */
int evp_parse_dd_keys(const char *path, evp_dd_keys *dd)
{
    struct {
        const char *name;
        double *p;
        int len;
    } keys[] = {
        {"a0", dd->a0, 2}, {"L0", &dd->L0, 1}, {"GND_ScaleRatio", &dd->GND_ScaleRatio, 1},
        {"T__K", &dd->T__K, 1}, {"Q_slip", dd->Q_slip, 2}, {"Q_bulk", dd->Q_bulk, 2},
        {"C_1", dd->C[0], 2}, {"C_2", dd->C[1], 2}, {"C_3", dd->C[2], 2}, {"C_4", dd->C[3], 2},
        {"C_5", dd->C[4], 2}, {"C_6", dd->C[5], 2}, {"C_7", dd->C[6], 2}, {"C_8", dd->C[7], 2},
        {"Selfinter", dd->inter[0], 2}, {"Coplanar", dd->inter[1], 2},
        {"CrossSlip", dd->inter[2], 2}, {"GlissileJunction", dd->inter[3], 2},
        {"HirthLock", dd->inter[4], 2}, {"LomerCottrellLock", dd->inter[5], 2},
        {"rho_SSD_initial", &dd->rho_SSD_initial, 1}, {"D_0", &dd->D_0, 1},
        {"tau_DRX", &dd->tau_DRX, 1}, {"DeltaTau_DRX", &dd->DeltaTau_DRX, 1},
    };
    const int nk = (int)(sizeof(keys) / sizeof(keys[0]));
    int seen[sizeof(keys) / sizeof(keys[0])] = {0};
    memset(dd, 0, sizeof(*dd));
    char buff[200];
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    while (1) {
        if (!fgets(buff, 200, fp)) buff[0] = '\n';
        if (feof(fp)) break;
        if (buff[0] == '#' || buff[0] == '\n') continue;
        char *tok = strtok(buff, " \t\n");
        if (!tok) break;
        for (int k = 0; k < nk; k++) {
            if (strcmp(tok, keys[k].name) != 0 || seen[k]) continue;
            seen[k] = 1;
            for (int j = 0; j < keys[k].len; j++) {
                tok = strtok(NULL, ", \t\n");
                if (!tok) break;
                keys[k].p[j] = atof(tok);
            }
            break;
        }
    }
    fclose(fp);
    return 0;
}

/* src/init.c:826: the Hill average of the cubic Voigt and Reuss shear moduli */
double evp_dd_shear_modulus(const double ec[3])
{
    return 0.1 * (ec[0] - ec[1] + 3 * ec[2]) +
           (0.5) * (5 * (ec[0] - ec[1]) * ec[2]) / (4 * ec[2] + 3 * (ec[0] - ec[1]));
}

/* Arsenlis-Parks interaction classes of the 12 FCC systems in the order of
 * the reference's .sx files, one row per system over all 12 columns:
 * S self, C coplanar, X cross slip, G glissile junction, H Hirth lock,
 * L Lomer-Cottrell lock.                                                    */
static const char *const FCC_CLASS[12] = {
    "SCCHGLXGGHLG", "CSCGXGGHLLHG", "CCSLGHGLHGGX", "HGLSCCHLGXGG",
    "GXGCSCLHGGHL", "LGHCCSGGXGLH", "XGGHLGSCCHGL", "GHLLHGCSCGXG",
    "GLHGGXCCSLGH", "HLGXGGHGLSCC", "LHGGHLGXGCSC", "GGXGLHLGHCCS",
};

void evp_dd_phase_setup(const evp_dd_keys *dd, int ph, const evpgpu_phase *phase, double shear_G,
                        double poisson_nu, evpgpu_dd_phase *out)
{
    memset(out, 0, sizeof(*out));
    for (int k = 0; k < 8; k++) out->c[k] = dd->C[k][ph];
    out->q_slip = dd->Q_slip[ph];
    out->q_bulk = dd->Q_bulk[ph];
    out->shear_G = shear_G;
    out->poisson_nu = poisson_nu;
    const int ns = phase->nsys;
    for (int j = 0; j < ns; j++) {
        const double *n = phase->bn[j], *b = phase->bb[j];
        out->bbB[j] = dd->a0[ph] / sqrt(2.0); /* FCC (DD_BCC off) */
        out->bp[j][0] = b[1] * n[2] - b[2] * n[1];
        out->bp[j][1] = b[2] * n[0] - b[0] * n[2];
        out->bp[j][2] = b[0] * n[1] - b[1] * n[0];
    }
    static const char codes[] = "SCXGHL";
    for (int i = 0; i < 12; i++)
        for (int j = 0; j < 12; j++)
            out->chi[i][j] = dd->inter[strchr(codes, FCC_CLASS[i][j]) - codes][ph];
    for (int i = 0; i < ns; i++)
        for (int j = 0; j < ns; j++) {
            const double *ni = phase->bn[i], *nj = phase->bn[j], *bj = phase->bb[j];
            double line[3] = {nj[1] * bj[2] - nj[2] * bj[1], nj[2] * bj[0] - nj[0] * bj[2],
                              nj[0] * bj[1] - nj[1] * bj[0]};
            double cs = fabs(ni[0] * line[0] + ni[1] * line[1] + ni[2] * line[2]);
            double cg1 = fabs(ni[0] * bj[0] + ni[1] * bj[1] + ni[2] * bj[2]);
            double cg3 = fabs(ni[0] * nj[0] + ni[1] * nj[1] + ni[2] * nj[2]);
            out->cth[0][i][j] = cs;
            out->cth[1][i][j] = cg1;
            out->cth[2][i][j] = cs;
            out->cth[3][i][j] = cg3;
            out->cth[4][i][j] = fabs(sqrt(fabs(1.0 - cs * cs)));
            out->cth[5][i][j] = fabs(sqrt(fabs(1.0 - cg1 * cg1)));
            out->cth[6][i][j] = out->cth[4][i][j];
            out->cth[7][i][j] = fabs(sqrt(fabs(1.0 - cg3 * cg3)));
        }
}
