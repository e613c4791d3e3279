/* evp_io.c -- input formats of the drop-in: the namelist (src/io.c:25-211),
 * the .sx single-crystal file (src/init.c:52-231) and the tensor helpers the
 * setup uses (src/kinematics.c:14-280, src/init.c:10-50).  The parsers keep
 * the reference's reading rules: 200-byte namelist lines, 80-byte .sx and
 * microstructure lines, fixed .sx line positions, first definition of a key
 * wins.                                                                      */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "evp_host.h"

/* ------------------------------------------------------------- namelist -- */
enum { NL_I, NL_R, NL_S };
typedef struct {
    const char *name;
    int type, len;
    size_t off;
} nl_key;

#define KI(f, n) {#f, NL_I, n, offsetof(evp_namelist, f)}
#define KR(f, n) {#f, NL_R, n, offsetof(evp_namelist, f)}
#define KS(f) {#f, NL_S, 1, offsetof(evp_namelist, f)}
#include <stddef.h>

static const nl_key KEYS[21] = {
    KI(CellDim, 3), KI(N_phases, 1), KI(Type_phases, 2), KR(TimeStep, 1), KI(N_steps, 1),
    KI(Update_Flag, 1), KI(Hard_Flag, 1), KI(Tex_Flag, 1), KI(IterMax, 1), KI(PrintControl, 2),
    KR(Err, 1), KI(VelGrad_BC_Flag, 6), KR(VelGrad_BC, 6), KI(Stress_BC_Flag, 6), KR(Stress_BC, 6),
    KI(ElastBC, 1), KR(ElastConst_PhaseI, 4), KR(ElastConst_PhaseII, 4), KS(Slip_PhaseI),
    KS(Slip_PhaseII), KS(initial_ms)};

/* GetNameList (src/io.c:96-164).  Returns the reference's ok flag, or -1 when
 * the file cannot be opened.                                                 */
int evp_parse_namelist(const char *path, evp_namelist *nl)
{
    char buff[200];
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    for (int k = 0; k < 21; k++) nl->status[k] = 0;
    int ok = 1;
    while (1) {
        if (!fgets(buff, 200, fp)) buff[0] = '\n';
        if (feof(fp)) break;
        if (buff[0] == '#' || buff[0] == '\n') continue;
        char *tok = strtok(buff, " \t\n");
        if (!tok) break; /* a blank line of spaces ends the namelist */
        int match = 0;
        for (int k = 0; k < 21; k++) {
            if (strcmp(tok, KEYS[k].name) != 0) continue;
            match = 1;
            if (nl->status[k] == 0) {
                nl->status[k] = 1;
                char *base = (char *)nl + KEYS[k].off;
                for (int j = 0; j < KEYS[k].len; j++) {
                    tok = strtok(NULL, ", \t\n");
                    if (tok) {
                        if (KEYS[k].type == NL_I) ((int *)base)[j] = (int)atol(tok);
                        else if (KEYS[k].type == NL_R) ((double *)base)[j] = atof(tok);
                        else {
                            strncpy(base, tok, 99);
                            base[99] = 0;
                        }
                    } else {
                        nl->status[k] = 2;
                        ok = 0;
                    }
                }
                tok = strtok(NULL, ", \t\n");
                if (tok) {
                    nl->status[k] = 3;
                    ok = 0;
                }
            } else {
                nl->status[k] = 4;
                ok = 0;
            }
            break;
        }
        if (!match) ok = 0;
    }
    fclose(fp);
    for (int k = 0; k < 21; k++)
        if (nl->status[k] != 1) ok = 1;
    return ok;
}

/* PrintNameList (src/io.c:166-211) */
void evp_print_namelist(const evp_namelist *nl)
{
    printf("===================================================\n");
    printf("-----------NameList for input data-----------------\n\n");
    for (int k = 0; k < 21; k++) {
        const char *base = (const char *)nl + KEYS[k].off;
        printf("%s\t", KEYS[k].name);
        if (strlen(KEYS[k].name) < 8) printf("\t");
        if (nl->status[k] > 0) {
            for (int j = 0; j < KEYS[k].len; j++) {
                if (KEYS[k].type == NL_I) printf("%d ", ((const int *)base)[j]);
                else if (KEYS[k].type == NL_R) printf("%#e ", ((const double *)base)[j]);
                else printf("%s ", base);
            }
        }
        switch (nl->status[k]) {
        case 0: printf("** no data"); break;
        case 2: printf("** missing data"); break;
        case 3: printf("** extra data"); break;
        case 4: printf("** multiply defined"); break;
        }
        printf("\n");
    }
    printf("\n-------------------End of NameList-----------------\n");
    printf("===================================================\n\n");
    fflush(stdout);
}

FILE *evp_open_rel(const char *dir, const char *name)
{
    FILE *fp = fopen(name, "r");
    if (!fp && name[0] != '/' && dir && dir[0]) {
        char path[1024];
        snprintf(path, sizeof(path), "%s/%s", dir, name);
        fp = fopen(path, "r");
    }
    return fp;
}

/* ------------------------------------------------------------ tensors -- */
static double B[6][3][3];
static int B_ready = 0;

static void basis(void)
{
    if (B_ready) return;
    const double r2 = 0.70710678118654744, r3 = 0.57735026918962584, r6 = 0.40824829046386304;
    memset(B, 0, sizeof(B));
    B[1][0][0] = -1. * r6; B[1][1][1] = -1. * r6; B[1][2][2] = 2. * r6;
    B[0][0][0] = -1. * r2; B[0][1][1] = r2;
    B[2][1][2] = r2; B[2][2][1] = r2;
    B[3][0][2] = r2; B[3][2][0] = r2;
    B[4][0][1] = r2; B[4][1][0] = r2;
    B[5][0][0] = r3; B[5][1][1] = r3; B[5][2][2] = r3;
    B_ready = 1;
}

void evp_t2_to_v6(const double t[3][3], double v[6])
{
    basis();
    for (int i = 0; i < 6; i++) {
        double s = 0.0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) s += t[a][b] * B[i][a][b];
        v[i] = s;
    }
}

void evp_v6_to_t2(const double v[6], double t[3][3])
{
    basis();
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double s = 0.0;
            for (int i = 0; i < 6; i++) s += v[i] * B[i][a][b];
            t[a][b] = s;
        }
}

void evp_t2_to_v5(const double t[3][3], double v[5])
{
    basis();
    for (int i = 0; i < 5; i++) {
        double s = 0.0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) s += t[a][b] * B[i][a][b];
        v[i] = s;
    }
}

void evp_v5_to_t2(const double v[5], double t[3][3])
{
    basis();
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double s = 0.0;
            for (int i = 0; i < 5; i++) s += v[i] * B[i][a][b];
            t[a][b] = s;
        }
}

void evp_c2_to_t4(const double c[6][6], double t4[3][3][3][3])
{
    basis();
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++) {
                    double s = 0.0;
                    for (int i = 0; i < 6; i++)
                        for (int j = 0; j < 6; j++) s += c[i][j] * B[i][a][b] * B[j][k][l];
                    t4[a][b][k][l] = s;
                }
}

void evp_t4_to_c2(const double t4[3][3][3][3], double c[6][6])
{
    basis();
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            double s = 0.0;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++)
                    for (int k = 0; k < 3; k++)
                        for (int l = 0; l < 3; l++) s += t4[a][b][k][l] * B[i][a][b] * B[j][k][l];
            c[i][j] = s;
        }
}

/* LU_inv_66 with LU_dcmp/LU_bksb rules (src/kinematics.c:282-428) */
int evp_lu_inv66(double a[6][6])
{
    double vv[6], y[6][6];
    int indx[6];
    for (int i = 0; i < 6; i++) {
        double big = 0.0, t;
        for (int j = 0; j < 6; j++)
            if ((t = fabs(a[i][j])) > big) big = t;
        if (fabs(big) < 1E-5) return EVPGPU_ERR_SINGULAR;
        vv[i] = 1.0 / big;
    }
    for (int j = 0; j < 6; j++) {
        int imax = j;
        for (int i = 0; i < j; i++) {
            double sum = a[i][j];
            for (int k = 0; k < i; k++) sum -= a[i][k] * a[k][j];
            a[i][j] = sum;
        }
        double big = 0.0;
        for (int i = j; i < 6; i++) {
            double sum = a[i][j];
            for (int k = 0; k < j; k++) sum -= a[i][k] * a[k][j];
            a[i][j] = sum;
            double dum = vv[i] * fabs(sum);
            if (dum >= big) {
                big = dum;
                imax = i;
            }
        }
        if (j != imax) {
            for (int k = 0; k < 6; k++) {
                double d = a[imax][k];
                a[imax][k] = a[j][k];
                a[j][k] = d;
            }
            vv[imax] = vv[j];
        }
        indx[j] = imax;
        if (fabs(a[j][j]) < 1E-5) a[j][j] = 1.0E-20;
        if (j != 5) {
            double dum = 1.0 / a[j][j];
            for (int i = j + 1; i < 6; i++) a[i][j] *= dum;
        }
    }
    for (int col = 0; col < 6; col++) {
        double b[6] = {0, 0, 0, 0, 0, 0};
        b[col] = 1.0;
        int ii = -1;
        for (int i = 0; i < 6; i++) {
            int ip = indx[i];
            double sum = b[ip];
            b[ip] = b[i];
            if (ii >= 0)
                for (int j = ii; j <= i - 1; j++) sum -= a[i][j] * b[j];
            else if (fabs(sum) > 1E-5)
                ii = i;
            b[i] = sum;
        }
        for (int i = 5; i >= 0; i--) {
            double sum = b[i];
            for (int j = i + 1; j < 6; j++) sum -= a[i][j] * b[j];
            b[i] = sum / a[i][i];
        }
        for (int i = 0; i < 6; i++) y[i][col] = b[i];
    }
    memcpy(a, y, sizeof(y));
    return 0;
}

/* EulerToTransMatrix opt 2 (src/kinematics.c:219-231) */
void evp_euler_to_mat(double t1, double ph, double t2, double a[3][3])
{
    double ct1 = cos(t1), st1 = sin(t1), cph = cos(ph), sph = sin(ph), ct2 = cos(t2), st2 = sin(t2);
    a[0][0] = ct1 * ct2 - st1 * st2 * cph;
    a[0][1] = st1 * ct2 + ct1 * st2 * cph;
    a[0][2] = st2 * sph;
    a[1][0] = -1.0 * ct1 * st2 - st1 * ct2 * cph;
    a[1][1] = -1.0 * st1 * st2 + ct1 * ct2 * cph;
    a[1][2] = ct2 * sph;
    a[2][0] = st1 * sph;
    a[2][1] = -1.0 * ct1 * sph;
    a[2][2] = cph;
}

/* Ten4thTransform opt 2 (src/kinematics.c:262-273) */
void evp_ten4_transform_inv(double c1[3][3][3][3], const double a[3][3], const double c2[3][3][3][3])
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++) {
                    double s = 0.0;
                    for (int m = 0; m < 3; m++)
                        for (int n = 0; n < 3; n++)
                            for (int p = 0; p < 3; p++)
                                for (int q = 0; q < 3; q++)
                                    s += a[m][i] * a[n][j] * a[p][k] * a[q][l] * c2[m][n][p][q];
                    c1[i][j][k][l] = s;
                }
}

/* ElastStiffnessMatrix (src/init.c:10-50) */
void evp_cubic_stiffness(const double *ec, double c[3][3][3][3])
{
    memset(c, 0, sizeof(double) * 81);
    for (int i = 0; i < 3; i++) {
        c[i][i][i][i] = ec[0];
        for (int j = 0; j < 3; j++)
            if (i != j) {
                c[i][i][j][j] = ec[1];
                c[i][j][i][j] = ec[2];
                c[i][j][j][i] = ec[2];
            }
    }
}

/* ---------------------------------------------------------------- .sx -- */
int evp_parse_sx(const char *dir, const char *name, evpgpu_phase *ph)
{
    char buf[80] = {0};
    FILE *fp = evp_open_rel(dir, name);
    if (!fp) return EVP_ERR_FILE;
    int nmax = 0, nmodes = 0, imode[16] = {0}, nsmx = 0, isectwx = 0;
    double nrsx = 0, gamd0x = 0, twshx = 0, t0f = 0, t0b = 0, t1 = 0, th0 = 0, th1 = 0, hs = 0, hl = 0;
    double xa[3];
    char xs[80];
#define RL() (void)fgets(buf, 80, fp)
    RL(); RL(); sscanf(buf, "%79s", xs);
    RL(); RL(); RL(); sscanf(buf, "%lf %lf %lf", &xa[0], &xa[1], &xa[2]);
    RL(); RL(); RL(); sscanf(buf, "%d", &nmax);
    RL(); RL(); RL(); sscanf(buf, "%d", &nmodes);
    if (nmodes > 16) nmodes = 16;
    RL(); RL();
    for (int i = 0; i < nmodes; i++) {
        RL();
        sscanf(buf, "%d", &imode[i]);
    }
    RL();
    for (int i = 0; i < nmax; i++) {
        int mx = -1, match = 0;
        RL();
        sscanf(buf, "%*s %*s %d %*s", &mx);
        for (int j = 0; j < nmodes; j++)
            if (mx == imode[j]) {
                match = 1;
                break;
            }
        RL(); RL(); sscanf(buf, "%d %lf %lf %lf %d", &nsmx, &nrsx, &gamd0x, &twshx, &isectwx);
        RL(); RL(); sscanf(buf, "%lf %lf %lf %lf %lf", &t0f, &t0b, &t1, &th0, &th1);
        RL(); RL(); sscanf(buf, "%lf %lf", &hs, &hl);
        RL();
        if (match) {
            if (nsmx > EVP_NSYSMX || nsmx < 0) {
                fclose(fp);
                return EVP_ERR_NSYS;
            }
            ph->nsys = nsmx;
            if (th0 < th1) {
                fclose(fp);
                return EVP_ERR_THET;
            }
            if (t1 < 1.E-6) {
                t1 = 1.E-6;
                th0 = th1;
            }
            for (int j = 0; j < nsmx; j++) {
                ph->nrs[j] = nrsx;
                ph->gam0[j] = gamd0x;
                ph->tau[j][0] = t0f;
                ph->tau[j][1] = t0b;
                ph->tau[j][2] = t1;
                ph->thet[j][0] = th0;
                ph->thet[j][1] = th1;
                RL();
                sscanf(buf, "%lf %lf %lf %lf %lf %lf", &ph->bn[j][0], &ph->bn[j][1], &ph->bn[j][2],
                       &ph->bb[j][0], &ph->bb[j][1], &ph->bb[j][2]);
            }
        } else {
            for (int j = 0; j < nsmx; j++) RL();
        }
    }
#undef RL
    fclose(fp);
    /* normalisation, orthogonality, Schmid tensors; repeated per used mode
       as the reference does (src/init.c:163-210) */
    for (int i = 0; i < nmodes; i++)
        for (int j = 0; j < ph->nsys; j++) {
            double *n = ph->bn[j], *b = ph->bb[j];
            double norm = sqrt(pow(n[0], 2.0) + pow(n[1], 2.0) + pow(n[2], 2.0));
            for (int m = 0; m < 3; m++) n[m] /= norm;
            norm = sqrt(pow(b[0], 2.0) + pow(b[1], 2.0) + pow(b[2], 2.0));
            for (int m = 0; m < 3; m++) b[m] /= norm;
            norm = 0.0;
            for (int m = 0; m < 3; m++) norm += n[m] * b[m];
            if (norm > 1E-3) return EVP_ERR_ORTHO;
            double a33[3][3];
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++) a33[a][c] = (n[a] * b[c] + n[c] * b[a]) / 2.0;
            evp_t2_to_v5(a33, ph->schm_xt[j]);
        }
    /* self/latent hardening from the last mode listed (Q9, src/init.c:219-228) */
    for (int i = 0; i < ph->nsys; i++)
        for (int j = 0; j < ph->nsys; j++) ph->hard[i][j] = (i == j) ? hs : hl;
    return 0;
}
