/*
 * evp_oracle.c -- CPU restatement of M-THEMES' EVPFFT Green-operator
 * fixed-point loop (power-law + Voce build, and the dislocation-density build
 * the reference ships; FP64).
 *
 * TEST INFRASTRUCTURE: the checker for the HIP path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Follows, in operation order:
 *   src/kinematics.c:14-484   basis change, rotations, LU, von Mises, Schmid
 *   src/constitutive.c:5-67   power-law strain rate (StrainRate_eval)
 *   src/constitutive.c:142-326 update_orient/Orientation, harden (Voce)
 *   src/evolution.c:2-310     NR_quality, UpdateStress (NR_MODIFIED line search)
 *   src/evolution.c:473-547   get_smacro
 *   src/evolution.c:1318-1547 the live increment (prologue, Green loop, epilogue)
 *   src/init.c:10-396,444-591,727-900,1135-1205  SetupJob pieces
 *   src/io.c:96-164           namelist parser
 *   dislocation-density law: src/constitutive.c:507-721 (StrainRate_Orowan),
 *   875-1108 (Gradient_ShearRate), 1346-1520 (DislocationEvolution);
 *   src/init.c:179-203, 593-702, 821-846, 1025, 1073-1125 (its set-up)
 *
 * Parity pins: tests/test_oracle_pins.py, tests/test_oracle_plastic.py
 * (the plastic whole loop against exact single-crystal and laminate
 * solutions), tests/test_oracle_line_search.py (every line-search trial),
 * tests/test_oracle_dd.py (the DD law; the SSD rate of dd_evolution one term
 * at a time against its physics, the GND rates by a closed form) and
 * tests/test_loopctl_step.py (get_smacro's core against the device's step).
 *
 * Dispositions of the reference's quirks (shared with the HIP path):
 *   Q1  the power-law driver is reconstructed: the live loop of
 *       src/evolution.c:1318-1547 with StrainRate_eval, then
 *       `if (Hard_Flag==1 && istep>1) harden();` after update_orient
 *       (src/evolution.c:1551,1571-1576).
 *   Q3  s_vm = sqrt(3/2 sum (dev_ij - delta_ij tr(dev)/3)^2) with the trace
 *       taken over the 3 valid diagonal entries (src/evolution.c:531 reads
 *       out of bounds).
 *   Q4  the displacement block (src/evolution.c:1578-1613) is not restated.
 *   Q10 |x|^n via pow(fabs(x), n): identical to fabs(pow(x, n)) for integer n,
 *       defined for non-integer n.
 *   Q11 harden(): when the slip increment is exactly 0 the CRSS is left
 *       unchanged (the reference computes 0*x/0 = NaN).
 *   Q12 Schmid tensors are initialised at setup.
 *   Q18 the cubic backtrack of the line search reads f2, which the reference
 *       never assigns (src/evolution.c:221-246, uninitialised); the restated
 *       algorithm sets f2 = f_new at each backtrack, as in the line search the
 *       code follows (Numerical Recipes lnsrch).
 *   Q19 StrainRate_Orowan reads rho_m for the passing stress before the same
 *       call recomputes it: the first call after the densities change uses
 *       the previous call's mobile density.  Reproduced.
 *   Q20 Gradient_ExchangeGrainID compares the z-minus side with the -y
 *       neighbour's grain (src/constitutive.c:818, 838, 858).  Reproduced.
 *   Q21 DislocationEvolution: the istep > 20000000 climb branch and the
 *       second-order GND terms (compiled out) are not restated.
 */
#include "evp_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_REF 3.14159265359      /* src/evp.h:77 (truncated, Q8) */
#define TWO_PI_REF 6.28318530718  /* src/evp.h:78 */
#define EPS_LU 1E-5               /* EPSILON_FFT, src/evp.h:87 */
#define TINY_LU 1.0E-20           /* src/kinematics.c:3 */
#define MAXD(a, b) ((a) > (b) ? (a) : (b))

/* ------------------------------------------------------------------ */
/* basis table B (src/kinematics.c:28-54)                              */
/* ------------------------------------------------------------------ */
static double Bt[6][3][3];
static int Bt_ready = 0;

static void basis_init(void)
{
    const double r2 = 0.70710678118654744;
    const double r3 = 0.57735026918962584;
    const double r6 = 0.40824829046386304;
    if (Bt_ready) return;
    memset(Bt, 0, sizeof(Bt));
    Bt[1][0][0] = -1. * r6; Bt[1][1][1] = -1. * r6; Bt[1][2][2] = 2. * r6;
    Bt[0][0][0] = -1. * r2; Bt[0][1][1] = r2;
    Bt[2][1][2] = r2; Bt[2][2][1] = r2;
    Bt[3][0][2] = r2; Bt[3][2][0] = r2;
    Bt[4][0][1] = r2; Bt[4][1][0] = r2;
    Bt[5][0][0] = r3; Bt[5][1][1] = r3; Bt[5][2][2] = r3;
    Bt_ready = 1;
}

void orc_basis(double B[6][3][3])
{
    basis_init();
    memcpy(B, Bt, sizeof(Bt));
}

void orc_v6_to_t2(const double v[6], double t[3][3])
{
    basis_init();
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double s = 0.0;
            for (int i = 0; i < 6; i++) s += v[i] * Bt[i][a][b];
            t[a][b] = s;
        }
}

void orc_t2_to_v6(const double t[3][3], double v[6])
{
    basis_init();
    for (int i = 0; i < 6; i++) {
        double s = 0.0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) s += t[a][b] * Bt[i][a][b];
        v[i] = s;
    }
}

void orc_v5_to_t2(const double v[5], double t[3][3])
{
    basis_init();
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double s = 0.0;
            for (int i = 0; i < 5; i++) s += v[i] * Bt[i][a][b];
            t[a][b] = s;
        }
}

void orc_t2_to_v5(const double t[3][3], double v[5])
{
    basis_init();
    for (int i = 0; i < 5; i++) {
        double s = 0.0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) s += t[a][b] * Bt[i][a][b];
        v[i] = s;
    }
}

void orc_c2_to_t4(const double c[6][6], double t4[3][3][3][3])
{
    basis_init();
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++) {
                    double s = 0.0;
                    for (int i = 0; i < 6; i++)
                        for (int j = 0; j < 6; j++)
                            s += c[i][j] * Bt[i][a][b] * Bt[j][k][l];
                    t4[a][b][k][l] = s;
                }
}

void orc_t4_to_c2(const double t4[3][3][3][3], double c[6][6])
{
    basis_init();
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            double s = 0.0;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++)
                    for (int k = 0; k < 3; k++)
                        for (int l = 0; l < 3; l++)
                            s += t4[a][b][k][l] * Bt[i][a][b] * Bt[j][k][l];
            c[i][j] = s;
        }
}

/* EulerToTransMatrix opt 2 (src/kinematics.c:219-231); angles in radians */
void orc_euler_to_mat(double t1, double ph, double t2, double a[3][3])
{
    double ct1 = cos(t1), st1 = sin(t1), cph = cos(ph), sph = sin(ph);
    double ct2 = cos(t2), st2 = sin(t2);
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

/* Ten4thTransform opt 2 (src/kinematics.c:262-273): c1 = a^T-rotated c2 */
void orc_ten4_transform_inv(double c1[3][3][3][3], const double a[3][3],
                            const double c2[3][3][3][3])
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

/* Crout LU with implicit pivoting, thresholds as src/kinematics.c:282-354 */
static int lu_dcmp6(double a[6][6], int indx[6])
{
    double vv[6];
    const int n = 6;
    for (int i = 0; i < n; i++) {
        double big = 0.0, t;
        for (int j = 0; j < n; j++)
            if ((t = fabs(a[i][j])) > big) big = t;
        if (fabs(big) < EPS_LU) return ORC_ERR_SINGULAR;
        vv[i] = 1.0 / big;
    }
    for (int j = 0; j < n; j++) {
        int imax = j;
        for (int i = 0; i < j; i++) {
            double sum = a[i][j];
            for (int k = 0; k < i; k++) sum -= a[i][k] * a[k][j];
            a[i][j] = sum;
        }
        double big = 0.0;
        for (int i = j; i < n; i++) {
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
            for (int k = 0; k < n; k++) {
                double dum = a[imax][k];
                a[imax][k] = a[j][k];
                a[j][k] = dum;
            }
            vv[imax] = vv[j];
        }
        indx[j] = imax;
        if (fabs(a[j][j]) < 1E-5) a[j][j] = TINY_LU;
        if (j != n - 1) {
            double dum = 1.0 / (a[j][j]);
            for (int i = j + 1; i < n; i++) a[i][j] *= dum;
        }
    }
    return ORC_OK;
}

/* src/kinematics.c:356-394; forward-substitution start threshold 1e-5 */
static void lu_bksb6(double a[6][6], const int indx[6], double b[6])
{
    const int n = 6;
    int ii = -1;
    for (int i = 0; i < n; i++) {
        int ip = indx[i];
        double sum = b[ip];
        b[ip] = b[i];
        if (ii >= 0)
            for (int j = ii; j <= i - 1; j++) sum -= a[i][j] * b[j];
        else if (fabs(sum) > 1E-5)
            ii = i;
        b[i] = sum;
    }
    for (int i = n - 1; i >= 0; i--) {
        double sum = b[i];
        for (int j = i + 1; j < n; j++) sum -= a[i][j] * b[j];
        b[i] = sum / a[i][i];
    }
}

/* LU_inv_66 (src/kinematics.c:396-428) */
int orc_lu_inv66(double c[6][6])
{
    int indx[6];
    double y[6][6], col[6];
    int st = lu_dcmp6(c, indx);
    if (st) return st;
    for (int j = 0; j < 6; j++) {
        for (int i = 0; i < 6; i++) col[i] = 0.0;
        col[j] = 1.0;
        lu_bksb6(c, indx, col);
        for (int i = 0; i < 6; i++) y[i][j] = col[i];
    }
    memcpy(c, y, sizeof(y));
    return ORC_OK;
}

/* VonMises (src/kinematics.c:465-484) */
double orc_vonmises(const double t[3][3])
{
    double trace = t[0][0] + t[1][1] + t[2][2];
    double vm = 0.0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double d = (t[a][b] + t[b][a]) / 2. - (double)(a == b) * trace / 3.0;
            vm += d * d;
        }
    return sqrt(2. / 3. * vm);
}

/* update_schmid body for one voxel (src/kinematics.c:430-463) */
void orc_rotate_schmid(const double T[3][3], int nsys, const double schm_xt[][5],
                       double schm_gr[][5])
{
    for (int s = 0; s < nsys; s++) {
        double a33[3][3], r33[3][3];
        orc_v5_to_t2(schm_xt[s], a33);
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                double acc = 0.0;
                for (int p = 0; p < 3; p++)
                    for (int q = 0; q < 3; q++) acc += T[a][p] * T[b][q] * a33[p][q];
                r33[a][b] = acc;
            }
        orc_t2_to_v5(r33, schm_gr[s]);
    }
}

/* Orientation (src/constitutive.c:142-194) */
void orc_orientation(double tran[3][3], double rot[3][3])
{
    double w[3] = {rot[2][1], rot[0][2], rot[1][0]};
    double wn = 0.0;
    for (int i = 0; i < 3; i++) wn += w[i] * w[i];
    wn = sqrt(wn);
    double wt = tan(wn / 2.0);
    if (fabs(wn) < 1E-6) wn = 1.0;
    for (int i = 0; i < 3; i++) w[i] *= wt / wn;
    double om[3][3] = {{0., -1. * w[2], w[1]}, {w[2], 0., -1. * w[0]}, {-1. * w[1], w[0], 0.}};
    double om2[3][3], tn[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double s = 0.0;
            for (int i = 0; i < 3; i++) s += om[a][i] * om[i][b];
            om2[a][b] = s;
        }
    wt = wt * wt;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            rot[a][b] = (double)(a == b) + 2. * (om[a][b] + om2[a][b]) / (1. + wt);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double s = 0.0;
            for (int i = 0; i < 3; i++) s += rot[a][i] * tran[i][b];
            tn[a][b] = s;
        }
    memcpy(tran, tn, sizeof(tn));
}

/* StrainRate_eval (src/constitutive.c:5-67), backstress xkin = 0
 * (src/init.c:1159).  Q10: |x|^n as pow(fabs(x), n).                  */
void orc_strain_rate(const double stress[6], double edot[6], double dedot[6][6],
                     int nsys, const double sc[][5], const double tau_t[][2],
                     const double nsr[], const double gam0[], double gamdot[])
{
    double rss1[ORC_NSYSMX], rss2[ORC_NSYSMX];
    for (int i = 0; i < nsys; i++) {
        double r = sc[i][0] * stress[0] + sc[i][1] * stress[1] + sc[i][2] * stress[2] +
                   sc[i][3] * stress[3] + sc[i][4] * stress[4];
        int isg = ((r - 0.0) < 0) ? 1 : 0;
        r = (r - 0.0) / tau_t[i][isg];
        rss1[i] = gam0[i] * nsr[i] * fabs(pow(fabs(r), nsr[i] - 1.)) / tau_t[i][isg];
        rss2[i] = gam0[i] * fabs(pow(fabs(r), nsr[i])) * ((double)(2 * (r > 0) - 1));
        gamdot[i] = rss2[i];
    }
    for (int i = 0; i < 5; i++) {
        double s = 0.;
        for (int j = 0; j < nsys; j++) s += sc[j][i] * rss2[j];
        edot[i] = s;
    }
    edot[5] = 0.0;
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
            double s = 0.0;
            for (int k = 0; k < nsys; k++) s += sc[k][i] * sc[k][j] * rss1[k];
            dedot[i][j] = s;
        }
    for (int i = 0; i < 6; i++) {
        dedot[i][5] = 0.0;
        dedot[5][i] = 0.0;
    }
}

/* harden() body for one voxel (src/constitutive.c:275-326), Q11 guard */
void orc_harden_voxel(int nsys, const double gamdot[], double dt,
                      const double hard[][ORC_NSYSMX], const double tau[][3],
                      const double thet[][2], double crss[][2], double *gamacum)
{
    double gamtot = *gamacum, deltgam = 0.0;
    for (int i = 0; i < nsys; i++) deltgam += fabs(gamdot[i]) * dt;
    for (int i = 0; i < nsys; i++) {
        double dtau = 0.;
        for (int j = 0; j < nsys; j++) dtau += hard[i][j] * fabs(gamdot[j]) * dt;
        double tau0 = tau[i][0], tau1 = tau[i][2];
        double th0 = thet[i][0], th1 = thet[i][1];
        double tiny = 1.E-4 * tau0, voce = 0.0;
        if (fabs(th0) > tiny) {
            voce = th1 * deltgam;
            if (fabs(tau1) > tiny) {
                double fact = fabs(th0 / tau1);
                double exp_ini = exp(-1.0 * gamtot * fact);
                double exp_del = exp(-1.0 * deltgam * fact);
                voce += -1. * (fact * tau1 - th1) / fact * exp_ini * (exp_del - 1.0) -
                        th1 / fact * exp_ini *
                            (exp_del * ((gamtot + deltgam) * fact + 1.0) - (gamtot * fact + 1.0));
            }
        }
        if (deltgam != 0.0) {
            crss[i][0] += dtau * voce / deltgam;
            crss[i][1] += dtau * voce / deltgam;
        }
    }
    *gamacum = gamtot + deltgam;
}

/* ------------------------------------------------------------------ */
/* dislocation-density law (DD_BASED_FLAG, DD_CU, FORWARD_PLUS_BACKWARD,
 * DD_GND_FIRST_ORDER, PF_DRX with no DRX event yet: src/evpFlags.h)    */
/* ------------------------------------------------------------------ */
#define TRIAL_FREQ_DRX_REF 1.0E10            /* src/evp.h:80 */
#define EV_IN_J_REF 1.60217733e-19           /* src/evp.h:81 */
#define J_IN_EV_REF (1.0 / EV_IN_J_REF)      /* src/evp.h:82 */
#define MY_J_IN_EV_REF 0.62415063631         /* src/evp.h:83 */
#define BOLZ_REF 1.380658e-23                /* src/evp.h:84-85 */
#define EPSILON_DD (5.0E-3)                  /* src/PFgrid.h:50, via src/pf_xrx.h:47 */

/* kT of src/evp.h:86 */
double orc_dd_kT(double T_K) { return (T_K * BOLZ_REF * J_IN_EV_REF); }

/* StrainRate_Orowan (src/constitutive.c:507-721), FORWARD_PLUS_BACKWARD
 * branch.  rho_m is read (passing stress) before it is recomputed, system by
 * system, so the first call after the densities change uses the previous
 * call's mobile density (Q19).  Returns ORC_ERR_NAN where the reference's
 * assert(isnan(...)==0) aborts.                                         */
int orc_strain_rate_dd(const double stress[6], double edot[6], double dedot[6][6], int nsys,
                       const double sc[][5], const orc_dd_phase *ph, const double rho_s[],
                       const double rho_g1[], const double rho_g2[], const double rho_g3[],
                       double rho_m[], double rho_F[], double rho_P[], double gamdot[])
{
    const double tau_0 = 1000;
    const double kT = ph->kT;
    double gam[ORC_NSYSMX], gam2[ORC_NSYSMX];
    const double (*cs)[ORC_NSYSMX] = ph->cth[0], (*cg1)[ORC_NSYSMX] = ph->cth[1];
    const double (*cg2)[ORC_NSYSMX] = ph->cth[2], (*cg3)[ORC_NSYSMX] = ph->cth[3];
    const double (*ss)[ORC_NSYSMX] = ph->cth[4], (*sg1)[ORC_NSYSMX] = ph->cth[5];
    const double (*sg2)[ORC_NSYSMX] = ph->cth[6], (*sg3)[ORC_NSYSMX] = ph->cth[7];
    for (int i = 0; i < nsys; i++) {
        rho_F[i] = 0.0;
        rho_P[i] = 0.0;
        for (int j = 0; j < nsys; j++) {
            rho_F[i] += ph->chi[i][j] * (rho_s[j] * cs[i][j] + fabs(rho_g1[j] * cg1[i][j]) +
                                         fabs(rho_g2[j] * cg2[i][j]) + fabs(rho_g3[j] * cg3[i][j]));
            rho_P[i] += ph->chi[i][j] * (rho_s[j] * ss[i][j] + fabs(rho_g1[j] * sg1[i][j]) +
                                         fabs(rho_g2[j] * sg2[i][j]) + fabs(rho_g3[j] * sg3[i][j]));
        }
    }
    for (int i = 0; i < nsys; i++) {
        const double b = ph->bbB[i];
        double tau_pass = ph->c1 * ph->G * b * sqrt(fabs(rho_P[i] + rho_m[i])) * 1.0E-3;
        tau_pass += ph->drx;
        double rss = sc[i][0] * stress[0] + sc[i][1] * stress[1] + sc[i][2] * stress[2] +
                     sc[i][3] * stress[3] + sc[i][4] * stress[4];
        /* the reference leaves tau_eff unset below the passing stress, where
         * it is not used */
        double tau_eff = 0.0;
        if (fabs(rss) > tau_pass) tau_eff = (fabs(rss) - tau_pass);
        double tau_cut = (kT / MY_J_IN_EV_REF) * sqrt(rho_F[i]) / ph->c2 / ph->c3 / b / b * 0.1;
        rho_m[i] = 2.0 * kT / ph->c1 / ph->c2 / ph->c3 / ph->G / (pow(b, 3.0)) *
                   sqrt(rho_P[i] * rho_F[i]) * 160.2177;
        double gam_o = 1E-3 * (rho_m[i]) * TRIAL_FREQ_DRX_REF * ph->c2 * b * exp(-1.0 * ph->q_slip) /
                       sqrt(rho_F[i]);
        double t = (fabs(rss) > tau_pass) ? gam_o * sinh(tau_eff / (tau_0 + tau_cut)) : 0.0;
        if (isnan(t)) return ORC_ERR_NAN;
        gam[i] = t * ((double)(2 * (rss > 0) - 1));
        t = (fabs(rss) > tau_pass) ? gam_o * cosh(tau_eff / (tau_cut + tau_0)) / (tau_cut + tau_0)
                                   : 0.0;
        if (isnan(t)) return ORC_ERR_NAN;
        gam2[i] = t * (fabs(rss) > tau_pass);
        gamdot[i] = gam[i];
    }
    for (int i = 0; i < 5; i++) {
        double s = 0.;
        for (int j = 0; j < nsys; j++) s += sc[j][i] * gam[j];
        if (isnan(s)) return ORC_ERR_NAN;
        edot[i] = s;
    }
    edot[5] = 0.0;
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
            double s = 0.0;
            for (int k = 0; k < nsys; k++) s += sc[k][i] * sc[k][j] * gam2[k];
            dedot[i][j] = s;
        }
    for (int i = 0; i < 6; i++) {
        dedot[i][5] = 0.0;
        dedot[5][i] = 0.0;
    }
    return ORC_OK;
}

/* InteractionStrengthMatrix("FCC", ...) (src/init.c:593-702): the upper
 * triangle as interaction types, row i listing j = i..11 (S self, C
 * coplanar, X cross slip, G glissile junction, H Hirth lock, L
 * Lomer-Cottrell lock; Arsenlis and Parks, JMPS 50 (2002) 1979), then
 * mirrored.  The values come from the namelist in the order Selfinter,
 * Coplanar, CrossSlip, GlissileJunction, HirthLock, LomerCottrellLock.    */
void orc_dd_interaction(const double inter[6], double chi[12][12])
{
    static const char *rows[12] = {"SCCHGLXGGHLG", "SCGXGGHLLHG", "SLGHGLHGGX", "SCCHLGXGG",
                                   "SCLHGGHL",     "SGGXGLH",     "SCCHGL",     "SCGXG",
                                   "SLGH",         "SCC",         "SC",         "S"};
    static const char *codes = "SCXGHL";
    for (int i = 0; i < 12; i++)
        for (int k = 0; rows[i][k]; k++) {
            int t = (int)(strchr(codes, rows[i][k]) - codes);
            chi[i][i + k] = inter[t];
        }
    for (int i = 1; i < 12; i++)
        for (int j = 0; j < i; j++) chi[i][j] = chi[j][i];
}

/* projection tables (src/init.c:1108-1125): cos/sin of the angles between
 * plane normal i and, for system j, the screw line direction n x b (SSD,
 * GND type 2), b (GND type 1) and n (GND type 3).  Order: cos s, g1, g2, g3,
 * sin s, g1, g2, g3.                                                        */
void orc_dd_angles(int nsys, const double bn[][3], const double bb[][3], double cth[8][12][12])
{
    for (int i = 0; i < nsys; i++)
        for (int j = 0; j < nsys; j++) {
            double sv[3];
            sv[0] = bn[j][1] * bb[j][2] - bn[j][2] * bb[j][1];
            sv[1] = bn[j][2] * bb[j][0] - bn[j][0] * bb[j][2];
            sv[2] = bn[j][0] * bb[j][1] - bn[j][1] * bb[j][0];
            double c = fabs(bn[i][0] * sv[0] + bn[i][1] * sv[1] + bn[i][2] * sv[2]);
            cth[0][i][j] = c;
            cth[4][i][j] = fabs(sqrt(fabs(1.0 - c * c)));
            c = fabs(bn[i][0] * bb[j][0] + bn[i][1] * bb[j][1] + bn[i][2] * bb[j][2]);
            cth[1][i][j] = c;
            cth[5][i][j] = fabs(sqrt(fabs(1.0 - c * c)));
            cth[2][i][j] = cth[0][i][j];
            cth[6][i][j] = cth[4][i][j];
            c = fabs(bn[i][0] * bn[j][0] + bn[i][1] * bn[j][1] + bn[i][2] * bn[j][2]);
            cth[3][i][j] = c;
            cth[7][i][j] = fabs(sqrt(fabs(1.0 - c * c)));
        }
}

/* shear modulus of the DD build (src/init.c:826): the Hill average of the
 * cubic Voigt and Reuss bounds, as the reference writes it                 */
double orc_dd_shear_modulus(const double ec[3])
{
    return 0.1 * (ec[0] - ec[1] + 3 * ec[2]) +
           (0.5) * (5 * (ec[0] - ec[1]) * ec[2]) / (4 * ec[2] + 3 * (ec[0] - ec[1]));
}

/* ------------------------------------------------------------------ */
/* per-voxel stress update (src/evolution.c:2-41, 44-310)             */
/* ------------------------------------------------------------------ */
typedef struct {
    const double (*sg)[6];
    const double (*c0)[6];
    const double *xl, *eps, *strain;
    double dt;
    int nsys;
    const double (*sc)[5];
    const double (*crss)[2];
    const double *nsr, *gam0;
    double *gamdot;
    /* dislocation-density law: phase constants and this voxel's densities */
    const orc_dd_phase *dd;
    const double *rho_s, *rho_g1, *rho_g2, *rho_g3;
    double *rho_m, *rho_F, *rho_P;
    int status;
} nrctx;

/* the rate law of the build: StrainRate_Orowan with DD_BASED_FLAG, else
 * StrainRate_eval (src/evolution.c:10-18, 116-125)                      */
static void nr_rate(nrctx *c, const double sig[6], double edot[6], double dedot[6][6])
{
    if (c->dd) {
        int st = orc_strain_rate_dd(sig, edot, dedot, c->nsys, c->sc, c->dd, c->rho_s, c->rho_g1,
                                    c->rho_g2, c->rho_g3, c->rho_m, c->rho_F, c->rho_P, c->gamdot);
        if (st && !c->status) c->status = st;
    } else {
        orc_strain_rate(sig, edot, dedot, c->nsys, c->sc, c->crss, c->nsr, c->gam0, c->gamdot);
    }
}

static double nr_quality(nrctx *c, const double sig[6])
{
    double edotp[6], dd[6][6], tot[6], res[6];
    nr_rate(c, sig, edotp, dd);
    for (int i = 0; i < 6; i++) {
        tot[i] = c->eps[i] + edotp[i] * c->dt;
        for (int j = 0; j < 6; j++) tot[i] += c->sg[i][j] * sig[j];
    }
    for (int i = 0; i < 6; i++) {
        res[i] = sig[i] - c->xl[i];
        for (int j = 0; j < 6; j++) res[i] += c->c0[i][j] * (tot[j] - c->strain[j]);
    }
    double f = 0.0;
    for (int i = 0; i < 6; i++) f += res[i] * res[i];
    return f / 2.0;
}

/* Newton tolerances: 1e-10 on the relative step and TOLX = 1e-7 on the line
 * search (src/evolution.c:105-110, 166-170); adjustable for validation runs. */
static double g_nr_tol = 1E-10, g_tolx = 1.E-7;

void orc_set_newton_tol(double tol, double tolx)
{
    g_nr_tol = tol;
    g_tolx = tolx;
}

/* Optional record of the line search (tests of its branches): one row of
 * ORC_LS_ROW doubles per trial point, see evp_oracle.h.  Not thread-safe:
 * set it only around single-voxel calls.                                  */
static double *g_ls_trace = NULL;
static int g_ls_cap = 0, g_ls_n = 0;

void orc_ls_trace(double *buf, int cap_rows)
{
    g_ls_trace = buf;
    g_ls_cap = buf ? cap_rows : 0;
    g_ls_n = 0;
}

int orc_ls_trace_count(void) { return g_ls_n; }

static void ls_record(int iterl, double alam, double f_new, double f_old, double slope,
                      double alamin, int outcome, double tmplam, const double old6[6],
                      const double step[6], double test, double sum, int scaled)
{
    if (!g_ls_trace || g_ls_n >= g_ls_cap) return;
    double *r = g_ls_trace + (size_t)g_ls_n * ORC_LS_ROW;
    r[0] = iterl; r[1] = alam; r[2] = f_new; r[3] = f_old; r[4] = slope; r[5] = alamin;
    r[6] = outcome; r[7] = tmplam;
    for (int i = 0; i < 6; i++) {
        r[8 + i] = old6[i];
        r[14 + i] = step[i];
    }
    r[20] = test; r[21] = sum; r[22] = scaled; r[23] = 0.0;
    g_ls_n++;
}

/* The backtracking step of the NR_MODIFIED line search
 * (src/evolution.c:218-249) after a rejected trial at alam (merit f_new),
 * the previous rejected one at alam2 (f2), f(0) = f_old, f'(0) = slope:
 * returns tmplam (before the caller's 0.1*alam floor); *outcome is the
 * branch taken (2..6, the trace codes).                                   */
double orc_ls_next_lambda(double alam, double alam2, double f_new, double f2, double f_old,
                          double slope, int *outcome)
{
    double tmplam;
    int oc;
    if (fabs(alam - 1.0) < 1.E-4) {
        tmplam = -slope / (2.0 * (f_new - f_old - slope));
        oc = 2;
    } else {
        double rhs1 = f_new - f_old - alam * slope;
        double rhs2 = f2 - f_old - alam2 * slope;
        double ac = (rhs1 / (alam * alam) - rhs2 / (alam2 * alam2)) / (alam - alam2);
        double bc = (-alam2 * rhs1 / (alam * alam) + alam * rhs2 / (alam2 * alam2)) /
                    (alam - alam2);
        if (fabs(ac) < 1E-4) {
            tmplam = -slope / (2.0 * bc);
            oc = 3;
        } else {
            double disc = bc * bc - 3.0 * ac * slope;
            if (disc < 0.0) {
                tmplam = 0.5 * alam;
                oc = 4;
            } else if (bc <= 0.0) {
                tmplam = (-bc + sqrt(disc)) / (3.0 * ac);
                oc = 5;
            } else {
                tmplam = -slope / (bc + sqrt(disc));
                oc = 6;
            }
        }
        if (tmplam > 0.5 * alam) tmplam = 0.5 * alam;
    }
    if (outcome) *outcome = oc;
    return tmplam;
}

/* UpdateStress for one voxel; dd (may be NULL) selects the dislocation-
 * density law and carries its phase constants and the voxel's densities. */
static int update_stress(const double sg66[6][6], const double c066[6][6], double dt,
                         double sig[3][3], const double eps[3][3], const double disgrad[3][3],
                         int nsys, const double sc[][5], const double crss[][2],
                         const double nsr[], const double gam0[], double edot_out[3][3],
                         double gamdot[], double *err_s, double *err_e, int *n_iter,
                         const nrctx *ddv)
{
    double xl_t[3][3], st_t[3][3];
    double xl6[6], sig6[6], eps6[6], str6[6], old6[6];
    double edotp6[6], dd[6][6], tot[6], res[6], jinv[6][6], jac[6][6];
    double conv_s = 0.0, conv_e = 0.0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            xl_t[a][b] = sig[a][b];
            st_t[a][b] = (disgrad[a][b] + disgrad[b][a]) / 2.0;
        }
    orc_t2_to_v6(xl_t, xl6);
    orc_t2_to_v6(xl_t, sig6);
    orc_t2_to_v6(eps, eps6);
    orc_t2_to_v6(st_t, str6);
    double signorm = 0.0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) signorm += xl_t[a][b] * xl_t[a][b];
    signorm = sqrt(signorm);

    nrctx c = {sg66, c066, xl6, eps6, str6, dt, nsys, sc, crss, nsr, gam0, gamdot};
    if (ddv) {
        c.dd = ddv->dd;
        c.rho_s = ddv->rho_s;
        c.rho_g1 = ddv->rho_g1;
        c.rho_g2 = ddv->rho_g2;
        c.rho_g3 = ddv->rho_g3;
        c.rho_m = ddv->rho_m;
        c.rho_F = ddv->rho_F;
        c.rho_P = ddv->rho_P;
    }
    const double erroral = g_nr_tol;
    int iterl = 0;
    double erral = 10 * erroral;
    while ((iterl < 500) && (fabs(erral) > fabs(erroral))) {
        iterl++;
        for (int i = 0; i < 6; i++) old6[i] = sig6[i];
        nr_rate(&c, sig6, edotp6, dd);
        if (c.status) return c.status;
        for (int i = 0; i < 6; i++) {
            tot[i] = eps6[i] + edotp6[i] * dt;
            for (int j = 0; j < 6; j++) tot[i] += sg66[i][j] * sig6[j];
        }
        for (int i = 0; i < 6; i++) {
            res[i] = sig6[i] - xl6[i];
            for (int j = 0; j < 6; j++) res[i] += c066[i][j] * (tot[j] - str6[j]);
        }
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) {
                jinv[i][j] = (double)(i == j);
                for (int k = 0; k < 6; k++) jinv[i][j] += c066[i][k] * (sg66[k][j] + dd[k][j] * dt);
            }
        memcpy(jac, jinv, sizeof(jac));
        int st = orc_lu_inv66(jinv);
        if (st) return st;
        double f_old = 0.0;
        for (int i = 0; i < 6; i++) f_old += res[i] * res[i];
        f_old *= 0.5;
        double gradf[6], step[6];
        for (int i = 0; i < 6; i++) {
            gradf[i] = 0.0;
            for (int j = 0; j < 6; j++) gradf[i] += res[j] * jac[j][i];
        }
        for (int i = 0; i < 6; i++) {
            step[i] = 0.0;
            for (int j = 0; j < 6; j++) step[i] -= jinv[i][j] * res[j];
        }
        const double stpmax = 10.0, ALF = 1.E-4, TOLX = g_tolx;
        double sum = 0.0, slope = 0.0, test = 0.0;
        for (int i = 0; i < 6; i++) sum += step[i] * step[i];
        sum = sqrt(sum);
        const int scaled = sum > stpmax;
        if (sum > stpmax)
            for (int i = 0; i < 6; i++) step[i] *= stpmax / sum;
        for (int i = 0; i < 6; i++) slope += gradf[i] * step[i];
        if (slope >= 0) return ORC_ERR_LINESEARCH;
        for (int i = 0; i < 6; i++) {
            double t = fabs(step[i]) / MAXD(fabs(old6[i]), 1.0);
            if (t > test) test = t;
        }
        double alamin = TOLX / test, alam = 1.0, alam2 = 0.0, f2 = 0.0, tmplam = 0.0;
        while (1) {
            for (int i = 0; i < 6; i++) sig6[i] = old6[i] + alam * step[i];
            double f_new = nr_quality(&c, sig6);
            if (c.status) return c.status;
            if (alam < alamin) {
                ls_record(iterl, alam, f_new, f_old, slope, alamin, 1, 0.0, old6, step, test, sum,
                          scaled);
                for (int i = 0; i < 6; i++) sig6[i] = old6[i];
                break;
            } else if (f_new <= f_old + ALF * alam * slope) {
                ls_record(iterl, alam, f_new, f_old, slope, alamin, 0, 0.0, old6, step, test, sum,
                          scaled);
                break;
            } else {
                int outcome;
                tmplam = orc_ls_next_lambda(alam, alam2, f_new, f2, f_old, slope, &outcome);
                ls_record(iterl, alam, f_new, f_old, slope, alamin, outcome, tmplam, old6, step,
                          test, sum, scaled);
            }
            alam2 = alam;
            f2 = f_new; /* Q18 */
            alam = MAXD(tmplam, 0.1 * alam);
        }
        double conv_nr = 0.0;
        conv_e = 0.0;
        conv_s = 0.0;
        for (int i = 0; i < 6; i++) {
            conv_nr += (sig6[i] - old6[i]) * (sig6[i] - old6[i]);
            conv_e += (tot[i] - str6[i]) * (tot[i] - str6[i]);
            conv_s += (sig6[i] - xl6[i]) * (sig6[i] - xl6[i]);
        }
        erral = sqrt(conv_nr) / signorm;
        conv_s = sqrt(conv_s);
        conv_e = sqrt(conv_e);
    }
    orc_v6_to_t2(sig6, sig);
    orc_v6_to_t2(edotp6, edot_out);
    *err_s = conv_s;
    *err_e = conv_e;
    *n_iter = iterl;
    return ORC_OK;
}

int orc_update_stress_voxel(const double sg66[6][6], const double c066[6][6], double dt,
                            double sig[3][3], const double eps[3][3],
                            const double disgrad[3][3], int nsys, const double sc[][5],
                            const double crss[][2], const double nsr[], const double gam0[],
                            double edot_out[3][3], double gamdot[], double *err_s,
                            double *err_e, int *n_iter)
{
    return update_stress(sg66, c066, dt, sig, eps, disgrad, nsys, sc, crss, nsr, gam0, edot_out,
                         gamdot, err_s, err_e, n_iter, NULL);
}

/* ------------------------------------------------------------------ */
/* FFT: unnormalised complex DFT, exponent sign `sign` (FFTW convention) */
/* ------------------------------------------------------------------ */
static void fft1d(double *x /* 2n */, int n, int sign, double *tmp /* 2n */)
{
    if ((n & (n - 1)) == 0) {
        /* iterative radix-2, decimation in time */
        for (int i = 1, j = 0; i < n; i++) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                double tr = x[2 * i], ti = x[2 * i + 1];
                x[2 * i] = x[2 * j];
                x[2 * i + 1] = x[2 * j + 1];
                x[2 * j] = tr;
                x[2 * j + 1] = ti;
            }
        }
        for (int len = 2; len <= n; len <<= 1) {
            int half = len >> 1;
            for (int k = 0; k < half; k++) {
                double ang = 2.0 * M_PI * (double)k / (double)len;
                double wr = cos(ang), wi = sign * sin(ang);
                for (int s = 0; s < n; s += len) {
                    double *u = x + 2 * (s + k), *v = x + 2 * (s + k + half);
                    double vr = v[0] * wr - v[1] * wi, vi = v[0] * wi + v[1] * wr;
                    v[0] = u[0] - vr;
                    v[1] = u[1] - vi;
                    u[0] = u[0] + vr;
                    u[1] = u[1] + vi;
                }
            }
        }
    } else {
        for (int k = 0; k < n; k++) {
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < n; j++) {
                long m = ((long)j * k) % n;
                double ang = 2.0 * M_PI * (double)m / (double)n;
                double wr = cos(ang), wi = sign * sin(ang);
                sr += x[2 * j] * wr - x[2 * j + 1] * wi;
                si += x[2 * j] * wi + x[2 * j + 1] * wr;
            }
            tmp[2 * k] = sr;
            tmp[2 * k + 1] = si;
        }
        memcpy(x, tmp, sizeof(double) * 2 * n);
    }
}

void orc_fft3d(double *d, int nx, int ny, int nz, int sign)
{
    int dims[3] = {nx, ny, nz};
    long strides[3] = {(long)ny * nz, nz, 1};
    for (int ax = 0; ax < 3; ax++) {
        int n = dims[ax];
        long st = strides[ax];
        int o1 = (ax == 0) ? ny : nx, o2 = (ax == 2) ? ny : nz;
        long s1 = (ax == 0) ? strides[1] : strides[0];
        long s2 = (ax == 2) ? strides[1] : strides[2];
#pragma omp parallel
        {
            double *line = (double *)malloc(sizeof(double) * 4 * n);
            double *tmp = line + 2 * n;
#pragma omp for collapse(2) schedule(static)
            for (int a = 0; a < o1; a++)
                for (int b = 0; b < o2; b++) {
                    long base = a * s1 + b * s2;
                    for (int i = 0; i < n; i++) {
                        line[2 * i] = d[2 * (base + i * st)];
                        line[2 * i + 1] = d[2 * (base + i * st) + 1];
                    }
                    fft1d(line, n, sign, tmp);
                    for (int i = 0; i < n; i++) {
                        d[2 * (base + i * st)] = line[2 * i];
                        d[2 * (base + i * st) + 1] = line[2 * i + 1];
                    }
                }
            free(line);
        }
    }
}

/* ------------------------------------------------------------------ */
/* job state                                                          */
/* ------------------------------------------------------------------ */
typedef double t2[3][3];

struct orc_job {
    /* namelist (src/io.c:25-94, power-law subset) */
    int CellDim[3], N_phases, Type_phases[2], N_steps, Update_Flag, Hard_Flag, Tex_Flag;
    int IterMax, PrintControl[2], VelGrad_BC_Flag[6], Stress_BC_Flag[6], ElastBC;
    double TimeStep, Err, VelGrad_BC[6], Stress_BC[6], ElastConst[2][4];
    char Slip[2][256], initial_ms[256], dir[512];
    /* material per phase (src/init.c:52-231) */
    int nSYS[2];
    double nSRS[2][12], gam0[2][12], tau[2][12][3], thet[2][12][2], Hard[2][12][12];
    double bn[2][12][3], bb[2][12][3], Schm_xt[2][12][5], Cijkl[2][3][3][3][3];
    /* grid */
    int Nxyz;
    double WGT, Wgt_ph1;
    int *grain, *phase;
    double (*C_gr)[6][6];
    double (*sg)[6][6]; /* LU_inv_66(C_gr), a per-voxel constant */
    t2 *TranMat, *Sig, *Eps, *Edot, *DisGrad, *VelGrad;
    double (*Schm_gr)[12][5], (*crss)[12][2], (*gamdot)[12], *gamacum;
    double *kbuf;         /* complex work [N^3] */
    t2 *ksr, *ksi;        /* k-space components */
    double (*g)[3];       /* unit wave vectors */
    double *err_s_v, *err_e_v;
    int *nr_v;
    /* macro */
    double C0[3][3][3][3], S0[3][3][3][3], C066[6][6], S066[6][6];
    t2 Udot, Udot_s, Udot_a, Scauchy, DisGradAvg, DisGradAvg_t, DisGradAvg_actual;
    t2 dDisGradAvg, dDisGradAvg_acum, SigAvg, SigAvg1, EpsAvg, EdotAvg, VelGradAvg;
    double d_vm, s_vm, e_vm, s_vm1, Err_e, Err_s, TimeTot, evmp, dvmp;
    int CREEP_FLAG, istep;
    double green_time, green_iters, newton_iters;
    int status;
    /* dislocation-density law: namelist (src/io.c:47-93), per-phase tables
     * (src/init.c:165-203, 1040-1125) and per-voxel densities (1E12/m^2)  */
    int law;
    double a0[2], L0, GND_ScaleRatio, T_K, Q_slip[2], Q_bulk[2], Ck[8][2], Inter[6][2];
    double rho_SSD_initial, D_0, tau_DRX, DeltaTau_DRX;
    double Shear_G[2], bb_B[2][12], bp[2][12][3], Chi[2][12][12], cth[2][8][12][12];
    double (*rho_s)[12], (*rho_m)[12], (*rho_g1)[12], (*rho_g2)[12], (*rho_g3)[12];
    double (*rho_F)[12], (*rho_P)[12], (*gadot_grad)[12][3], *t_last;
    double (*rho_dipole)[12], (*rho_forest)[12];
    double dd_avg[7]; /* AvgDislocation: F, P, s, m, |g|, dipole, forest */
    unsigned char (*gbc)[6]; /* GB_checkX/Y/Z [0] and [1] */
};

#define LOOPV for (int v = 0; v < J->Nxyz; v++)

static void voigt_to_full(const double v[6], t2 m)
{
    m[0][0] = v[0]; m[0][1] = v[5]; m[0][2] = v[4];
    m[1][0] = v[5]; m[1][1] = v[1]; m[1][2] = v[3];
    m[2][0] = v[4]; m[2][1] = v[3]; m[2][2] = v[2];
}

static FILE *open_rel(const orc_job *J, const char *name)
{
    FILE *fp = fopen(name, "r");
    if (!fp && name[0] != '/') {
        char path[1024];
        snprintf(path, sizeof(path), "%s/%s", J->dir, name);
        fp = fopen(path, "r");
    }
    return fp;
}

/* GetNameList (src/io.c:96-164): only the keys the power-law path reads */
static int parse_namelist(orc_job *J, const char *path)
{
    struct { const char *name; int type; void *ptr; int len; } tab[] = {
        {"CellDim", 0, J->CellDim, 3}, {"N_phases", 0, &J->N_phases, 1},
        {"Type_phases", 0, J->Type_phases, 2}, {"TimeStep", 1, &J->TimeStep, 1},
        {"N_steps", 0, &J->N_steps, 1}, {"Update_Flag", 0, &J->Update_Flag, 1},
        {"Hard_Flag", 0, &J->Hard_Flag, 1}, {"Tex_Flag", 0, &J->Tex_Flag, 1},
        {"IterMax", 0, &J->IterMax, 1}, {"PrintControl", 0, J->PrintControl, 2},
        {"Err", 1, &J->Err, 1}, {"VelGrad_BC_Flag", 0, J->VelGrad_BC_Flag, 6},
        {"VelGrad_BC", 1, J->VelGrad_BC, 6}, {"Stress_BC_Flag", 0, J->Stress_BC_Flag, 6},
        {"Stress_BC", 1, J->Stress_BC, 6}, {"ElastBC", 0, &J->ElastBC, 1},
        {"ElastConst_PhaseI", 1, J->ElastConst[0], 4},
        {"ElastConst_PhaseII", 1, J->ElastConst[1], 4},
        {"Slip_PhaseI", 2, J->Slip[0], 1}, {"Slip_PhaseII", 2, J->Slip[1], 1},
        {"initial_ms", 2, J->initial_ms, 1},
        /* DD_BASED_FLAG keys (src/io.c:47-68) and the two DRX keys the
         * passing stress reads (src/io.c:86-88) */
        {"a0", 1, J->a0, 2}, {"L0", 1, &J->L0, 1}, {"GND_ScaleRatio", 1, &J->GND_ScaleRatio, 1},
        {"T__K", 1, &J->T_K, 1}, {"Q_slip", 1, J->Q_slip, 2}, {"Q_bulk", 1, J->Q_bulk, 2},
        {"C_1", 1, J->Ck[0], 2}, {"C_2", 1, J->Ck[1], 2}, {"C_3", 1, J->Ck[2], 2},
        {"C_4", 1, J->Ck[3], 2}, {"C_5", 1, J->Ck[4], 2}, {"C_6", 1, J->Ck[5], 2},
        {"C_7", 1, J->Ck[6], 2}, {"C_8", 1, J->Ck[7], 2}, {"Selfinter", 1, J->Inter[0], 2},
        {"Coplanar", 1, J->Inter[1], 2}, {"CrossSlip", 1, J->Inter[2], 2},
        {"GlissileJunction", 1, J->Inter[3], 2}, {"HirthLock", 1, J->Inter[4], 2},
        {"LomerCottrellLock", 1, J->Inter[5], 2}, {"rho_SSD_initial", 1, &J->rho_SSD_initial, 1},
        {"D_0", 1, &J->D_0, 1}, {"tau_DRX", 1, &J->tau_DRX, 1},
        {"DeltaTau_DRX", 1, &J->DeltaTau_DRX, 1},
    };
    const int nt = (int)(sizeof(tab) / sizeof(tab[0]));
    int seen[64] = {0};
    char buff[200];
    FILE *fp = fopen(path, "r");
    if (!fp) return ORC_ERR_FILE;
    while (1) {
        if (!fgets(buff, 200, fp)) buff[0] = '\n';
        if (feof(fp)) break;
        if (buff[0] == '#' || buff[0] == '\n') continue;
        char *tok = strtok(buff, " \t\n");
        if (!tok) break;
        for (int k = 0; k < nt; k++) {
            if (strcmp(tok, tab[k].name) != 0) continue;
            if (!seen[k]) {
                seen[k] = 1;
                for (int j = 0; j < tab[k].len; j++) {
                    tok = strtok(NULL, ", \t\n");
                    if (!tok) break;
                    if (tab[k].type == 0) ((int *)tab[k].ptr)[j] = (int)atol(tok);
                    else if (tab[k].type == 1) ((double *)tab[k].ptr)[j] = atof(tok);
                    else strcpy((char *)tab[k].ptr, tok);
                }
            }
            break;
        }
    }
    fclose(fp);
    return ORC_OK;
}

/* ElastStiffnessMatrix (src/init.c:10-50): cubic c11, c12, c44 */
static void cubic_stiffness(const double *ec, double c[3][3][3][3])
{
    memset(c, 0, sizeof(double) * 81);
    double c11 = ec[0], c12 = ec[1], c44 = ec[2];
    for (int i = 0; i < 3; i++) {
        c[i][i][i][i] = c11;
        for (int j = 0; j < 3; j++)
            if (i != j) {
                c[i][i][j][j] = c12;
                c[i][j][i][j] = c44;
                c[i][j][j][i] = c44;
            }
    }
}

/* PlasticInit (src/init.c:52-231): fixed line positions, 80-char reads */
static int parse_sx(orc_job *J, const char *name, int ph)
{
    char buf[80] = {0};
    FILE *fp = open_rel(J, name);
    if (!fp) return ORC_ERR_FILE;
    int nmax = 0, nmodes = 0, imode[16] = {0};
    int nsmx = 0, isectwx = 0;
    double nrsx = 0, gamd0x = 0, twshx = 0, tau0xf = 0, tau0xb = 0, tau1x = 0, th0 = 0, th1 = 0;
    double hself = 0, hlat = 0, xa[3];
    char xs[80];
#define RL() (void)fgets(buf, 80, fp)
    RL(); RL(); sscanf(buf, "%79s", xs);
    RL(); RL(); RL(); sscanf(buf, "%lf %lf %lf", &xa[0], &xa[1], &xa[2]);
    RL(); RL(); RL(); sscanf(buf, "%d", &nmax);
    RL(); RL(); RL(); sscanf(buf, "%d", &nmodes);
    if (nmodes > 16) nmodes = 16;
    RL(); RL();
    for (int i = 0; i < nmodes; i++) { RL(); sscanf(buf, "%d", &imode[i]); }
    RL();
    for (int i = 0; i < nmax; i++) {
        int mx = -1, match = 0;
        RL(); sscanf(buf, "%*s %*s %d %*s", &mx);
        for (int j = 0; j < nmodes; j++)
            if (mx == imode[j]) { match = 1; break; }
        RL(); RL(); sscanf(buf, "%d %lf %lf %lf %d", &nsmx, &nrsx, &gamd0x, &twshx, &isectwx);
        RL(); RL(); sscanf(buf, "%lf %lf %lf %lf %lf", &tau0xf, &tau0xb, &tau1x, &th0, &th1);
        RL(); RL(); sscanf(buf, "%lf %lf", &hself, &hlat);
        RL();
        if (match) {
            if (nsmx > ORC_NSYSMX) { fclose(fp); return ORC_ERR_NSYS; }
            J->nSYS[ph] = nsmx;
            if (th0 < th1) { fclose(fp); return ORC_ERR_THET; }
            if (tau1x < 1.E-6) { tau1x = 1.E-6; th0 = th1; }
            for (int j = 0; j < nsmx; j++) {
                J->nSRS[ph][j] = nrsx;
                J->gam0[ph][j] = gamd0x;
                J->tau[ph][j][0] = tau0xf;
                J->tau[ph][j][1] = tau0xb;
                J->tau[ph][j][2] = tau1x;
                J->thet[ph][j][0] = th0;
                J->thet[ph][j][1] = th1;
                RL();
                sscanf(buf, "%lf %lf %lf %lf %lf %lf", &J->bn[ph][j][0], &J->bn[ph][j][1],
                       &J->bn[ph][j][2], &J->bb[ph][j][0], &J->bb[ph][j][1], &J->bb[ph][j][2]);
            }
        } else {
            for (int j = 0; j < nsmx; j++) RL();
        }
    }
#undef RL
    fclose(fp);
    for (int i = 0; i < nmodes; i++)
        for (int j = 0; j < J->nSYS[ph]; j++) {
            double *n = J->bn[ph][j], *b = J->bb[ph][j];
            double norm = sqrt(pow(n[0], 2.0) + pow(n[1], 2.0) + pow(n[2], 2.0));
            for (int m = 0; m < 3; m++) n[m] /= norm;
            norm = sqrt(pow(b[0], 2.0) + pow(b[1], 2.0) + pow(b[2], 2.0));
            for (int m = 0; m < 3; m++) b[m] /= norm;
            norm = 0.0;
            for (int m = 0; m < 3; m++) norm += n[m] * b[m];
            if (norm > 1E-3) return ORC_ERR_ORTHO;
            double a33[3][3];
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++) a33[a][c] = (n[a] * b[c] + n[c] * b[a]) / 2.0;
            orc_t2_to_v5(a33, J->Schm_xt[ph][j]);
        }
    for (int i = 0; i < J->nSYS[ph]; i++)
        for (int j = 0; j < J->nSYS[ph]; j++) J->Hard[ph][i][j] = (i == j) ? hself : hlat;
    return ORC_OK;
}

/* InitMicroStruct (src/init.c:233-396), single slab.  The file is read in
 * order; the per-voxel stiffness rotations then run in parallel and C066 is
 * summed in voxel order afterwards, as the reference's loop does.            */
static int read_ms(orc_job *J)
{
    char buf[80] = {0};
    FILE *fp = open_rel(J, J->initial_ms);
    if (!fp) return ORC_ERR_FILE;
    int nph1 = 0, Ny = J->CellDim[1], Nz = J->CellDim[2];
    double (*eul)[3] = calloc((size_t)J->Nxyz, sizeof(*eul));
    for (int v = 0; v < J->Nxyz; v++) {
        double ph = 0, th = 0, om = 0;
        int ii = 0, jj = 0, kk = 0, gr = 0, jp = 0;
        if (!fgets(buf, 80, fp)) buf[0] = 0;
        sscanf(buf, "%lf %lf %lf %d %d %d %d %d", &ph, &th, &om, &ii, &jj, &kk, &gr, &jp);
        int idx = ((ii - 1) * Ny + jj - 1) * Nz + kk - 1;
        if (idx != v || jp < 1 || jp > J->N_phases) {
            fclose(fp);
            free(eul);
            return ORC_ERR_MSINDEX;
        }
        if (jp == 1) nph1++;
        J->grain[v] = gr;
        J->phase[v] = jp;
        eul[v][0] = ph;
        eul[v][1] = th;
        eul[v][2] = om;
    }
    fclose(fp);
#pragma omp parallel for schedule(static)
    for (int v = 0; v < J->Nxyz; v++) {
        int jp = J->phase[v];
        if (J->Type_phases[jp - 1]) continue;
        double ph = eul[v][0], th = eul[v][1], om = eul[v][2];
        ph *= PI_REF / 180.;
        th *= PI_REF / 180.;
        om *= PI_REF / 180.;
        double sa2xt[3][3], c4[3][3][3][3];
        orc_euler_to_mat(ph, th, om, sa2xt);
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) J->TranMat[v][a][b] = sa2xt[b][a];
        orc_ten4_transform_inv(c4, sa2xt, J->Cijkl[jp - 1]);
        orc_t4_to_c2(c4, J->C_gr[v]);
    }
    free(eul);
    double c066l[6][6] = {{0}};
    for (int v = 0; v < J->Nxyz; v++) {
        if (J->Type_phases[J->phase[v] - 1]) continue;
        for (int a = 0; a < 6; a++)
            for (int b = 0; b < 6; b++) c066l[a][b] += J->C_gr[v][a][b];
    }
    J->Wgt_ph1 = (1.0 * (double)nph1) / J->Nxyz;
    for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++) J->C066[a][b] = c066l[a][b] * J->WGT;
    memcpy(J->S066, J->C066, sizeof(J->C066));
    int st = orc_lu_inv66(J->S066);
    if (st) return st;
    orc_c2_to_t4(J->C066, J->C0);
    orc_c2_to_t4(J->S066, J->S0);
    return ORC_OK;
}

#define ALLOC(p, n) (p) = calloc((size_t)(n), sizeof(*(p)))

static int g_law = ORC_LAW_POWER;
void orc_set_law(int law) { g_law = law; }
static void dd_grain_checks(orc_job *J);

/* the DD build's per-phase slip data (src/init.c:179-203): Burgers length
 * a0/sqrt(2) (FCC) and the sense vector p = b x n of the normalised axes   */
static void dd_phase_slip(orc_job *J, int ph)
{
    for (int j = 0; j < J->nSYS[ph]; j++) {
        const double *n = J->bn[ph][j], *b = J->bb[ph][j];
        J->bb_B[ph][j] = J->a0[ph] / sqrt(2.0);
        J->bp[ph][j][0] = b[1] * n[2] - b[2] * n[1];
        J->bp[ph][j][1] = b[2] * n[0] - b[0] * n[2];
        J->bp[ph][j][2] = b[0] * n[1] - b[1] * n[0];
    }
}

orc_job *orc_job_setup(const char *input_path, int *status)
{
    orc_job *J = calloc(1, sizeof(orc_job));
    int st;
    basis_init();
    strncpy(J->dir, input_path, sizeof(J->dir) - 1);
    char *sl = strrchr(J->dir, '/');
    if (sl) *sl = 0; else strcpy(J->dir, ".");
    if ((st = parse_namelist(J, input_path))) goto fail;
    J->Nxyz = J->CellDim[0] * J->CellDim[1] * J->CellDim[2];
    J->WGT = 1. / J->Nxyz;
    int N = J->Nxyz;
    ALLOC(J->grain, N); ALLOC(J->phase, N); ALLOC(J->C_gr, N); ALLOC(J->sg, N);
    ALLOC(J->TranMat, N); ALLOC(J->Sig, N); ALLOC(J->Eps, N); ALLOC(J->Edot, N);
    ALLOC(J->DisGrad, N); ALLOC(J->VelGrad, N); ALLOC(J->Schm_gr, N); ALLOC(J->crss, N);
    ALLOC(J->gamdot, N); ALLOC(J->gamacum, N); ALLOC(J->kbuf, 2 * (size_t)N);
    ALLOC(J->ksr, N); ALLOC(J->ksi, N); ALLOC(J->g, N);
    ALLOC(J->err_s_v, N); ALLOC(J->err_e_v, N); ALLOC(J->nr_v, N);

    /* k-space (src/init.c:775-796) */
    {
        int Nx = J->CellDim[0], Ny = J->CellDim[1], Nz = J->CellDim[2];
        for (int x = 0; x < Nx; x++)
            for (int y = 0; y < Ny; y++)
                for (int z = 0; z < Nz; z++) {
                    int v = (x * Ny + y) * Nz + z;
                    double t[3] = {(double)x * (double)(1.0 / Nx), (double)y * (double)(1.0 / Ny),
                                   (double)z * (double)(1.0 / Nz)};
                    for (int m = 0; m < 3; m++) {
                        t[m] -= round(t[m]);
                        t[m] *= TWO_PI_REF;
                    }
                    double g2 = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
                    if (x + y + z == 0) {
                        g2 = 1.0;
                        t[0] = t[1] = t[2] = 0.0;
                    }
                    double s = 1. / g2;
                    J->g[v][0] = t[0] * s;
                    J->g[v][1] = t[1] * s;
                    J->g[v][2] = t[2] * s;
                }
    }
    /* phases (src/init.c:815-862).  The DD build (DD_CU) replaces phase I's
     * constants by copper's and zeroes phase II's; both shear moduli come
     * from phase I (src/init.c:821-826, 840-845).                          */
    J->law = g_law;
    if (!J->Type_phases[0]) {
        if (J->law == ORC_LAW_DD) {
            J->ElastConst[0][0] = 241.1674 * 1000;
            J->ElastConst[0][1] = 89.1989 * 1000;
            J->ElastConst[0][2] = 151.9685 * 1000;
            J->Shear_G[0] = orc_dd_shear_modulus(J->ElastConst[0]);
        }
        cubic_stiffness(J->ElastConst[0], J->Cijkl[0]);
        if ((st = parse_sx(J, J->Slip[0], 0))) goto fail;
        if (J->law == ORC_LAW_DD) dd_phase_slip(J, 0);
    }
    if (J->N_phases == 2 && !J->Type_phases[1]) {
        if (J->law == ORC_LAW_DD) {
            J->ElastConst[1][0] = J->ElastConst[1][1] = J->ElastConst[1][2] = 0.0;
            J->Shear_G[1] = orc_dd_shear_modulus(J->ElastConst[0]);
        }
        cubic_stiffness(J->ElastConst[1], J->Cijkl[1]);
        if ((st = parse_sx(J, J->Slip[1], 1))) goto fail;
        if (J->law == ORC_LAW_DD) dd_phase_slip(J, 1);
    }
    if ((st = read_ms(J))) goto fail;
    if (J->law == ORC_LAW_DD) {
        /* densities, interaction matrix, projection tables and grain
         * checks (src/init.c:1025, 1073-1128) */
        ALLOC(J->rho_s, N); ALLOC(J->rho_m, N); ALLOC(J->rho_g1, N); ALLOC(J->rho_g2, N);
        ALLOC(J->rho_g3, N); ALLOC(J->rho_F, N); ALLOC(J->rho_P, N); ALLOC(J->gadot_grad, N);
        ALLOC(J->t_last, N); ALLOC(J->gbc, N); ALLOC(J->rho_dipole, N); ALLOC(J->rho_forest, N);
        for (int v = 0; v < N; v++) {
            J->t_last[v] = -1E3 * J->tau_DRX * J->TimeStep;
            for (int i = 0; i < ORC_NSYSMX; i++) J->rho_s[v][i] = J->rho_SSD_initial;
        }
        for (int ph = 0; ph < ORC_NPHMAX; ph++) {
            double in6[6];
            for (int t = 0; t < 6; t++) in6[t] = J->Inter[t][ph];
            orc_dd_interaction(in6, J->Chi[ph]);
            orc_dd_angles(J->nSYS[ph], (const double(*)[3])J->bn[ph], (const double(*)[3])J->bb[ph],
                          J->cth[ph]);
        }
        dd_grain_checks(J);
    }

    /* boundary conditions (src/init.c:864-900) */
    if ((J->VelGrad_BC_Flag[0] + J->VelGrad_BC_Flag[1] + J->VelGrad_BC_Flag[2]) == 2) {
        st = ORC_ERR_BC116;
        goto fail;
    }
    for (int i = 0; i < 6; i++) J->VelGrad_BC[i] *= J->VelGrad_BC_Flag[i];
    voigt_to_full(J->VelGrad_BC, J->Udot);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            J->Udot_s[a][b] = (J->Udot[a][b] + J->Udot[b][a]) / 2.;
            J->Udot_a[a][b] = (J->Udot[a][b] - J->Udot[b][a]) / 2.;
        }
    {
        double d6[6], d5[5], dev[3][3];
        orc_t2_to_v6(J->Udot_s, d6);
        for (int i = 0; i < 5; i++) d5[i] = d6[i];
        orc_v5_to_t2(d5, dev);
        J->d_vm = 0.0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) J->d_vm += pow(dev[a][b], 2.0);
        J->d_vm = sqrt(2. / 3. * J->d_vm);
    }
    for (int i = 0; i < 6; i++)
        if ((J->Stress_BC_Flag[i] * J->VelGrad_BC_Flag[i] != 0) ||
            ((J->Stress_BC_Flag[i] + J->VelGrad_BC_Flag[i]) != 1)) {
            st = ORC_ERR_BC114;
            goto fail;
        }
    J->CREEP_FLAG = 0;
    for (int i = 0; i < 6; i++) J->CREEP_FLAG += J->Stress_BC_Flag[i];
    J->CREEP_FLAG /= 6;
    voigt_to_full(J->Stress_BC, J->Scauchy);

    /* CRSS (src/init.c:1135-1163); sg = LU_inv_66(C_gr) is loop-invariant */
    int st_sg = ORC_OK;
#pragma omp parallel for schedule(static)
    for (int v = 0; v < N; v++) {
        int jp = J->phase[v];
        J->gamacum[v] = 0.0;
        if (J->Type_phases[jp - 1] == 0) {
            for (int i = 0; i < J->nSYS[jp - 1]; i++) {
                J->crss[v][i][0] = J->tau[jp - 1][i][0];
                J->crss[v][i][1] = J->tau[jp - 1][i][1];
            }
            memcpy(J->sg[v], J->C_gr[v], sizeof(J->sg[v]));
            int s2 = orc_lu_inv66(J->sg[v]);
            if (s2) {
#pragma omp critical
                st_sg = s2;
            }
            orc_rotate_schmid(J->TranMat[v], J->nSYS[jp - 1], J->Schm_xt[jp - 1], J->Schm_gr[v]);
        }
    }
    if ((st = st_sg)) goto fail;
    /* InitFields (src/init.c:444-515) */
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) J->DisGradAvg[a][b] = J->Udot[a][b] * J->TimeStep;
    for (int v = 0; v < N; v++) {
        int jp = J->phase[v];
        if (!J->Type_phases[jp - 1]) {
            double cg[3][3][3][3];
            orc_c2_to_t4(J->C_gr[v], cg);
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    double s = 0.0;
                    for (int p = 0; p < 3; p++)
                        for (int q = 0; q < 3; q++) s += cg[a][b][p][q] * J->DisGradAvg[p][q];
                    if (J->CREEP_FLAG == 1) s += J->Scauchy[a][b];
                    J->Sig[v][a][b] = s;
                }
        }
    }
    J->e_vm = J->d_vm * J->TimeStep;
    J->istep = 0;
    if (status) *status = ORC_OK;
    return J;
fail:
    if (status) *status = st;
    orc_job_free(J);
    return NULL;
}

void orc_job_free(orc_job *J)
{
    if (!J) return;
    void *ptrs[] = {J->grain, J->phase, J->C_gr, J->sg, J->TranMat, J->Sig, J->Eps, J->Edot,
                    J->DisGrad, J->VelGrad, J->Schm_gr, J->crss, J->gamdot, J->gamacum,
                    J->kbuf, J->ksr, J->ksi, J->g, J->err_s_v, J->err_e_v, J->nr_v,
                    J->rho_s, J->rho_m, J->rho_g1, J->rho_g2, J->rho_g3, J->rho_F, J->rho_P,
                    J->gadot_grad, J->t_last, J->gbc, J->rho_dipole, J->rho_forest};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    free(J);
}

/* M3Inverse (src/init.c:517-539) */
static void m3inv(double a[3][3], double ia[3][3])
{
    double deta = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                  a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                  a[0][2] * (a[1][0] * a[2][1] - a[2][0] * a[1][1]);
    ia[0][0] = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    ia[0][1] = a[0][2] * a[2][1] - a[0][1] * a[2][2];
    ia[0][2] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
    ia[1][0] = a[1][2] * a[2][0] - a[1][0] * a[2][2];
    ia[1][1] = a[0][0] * a[2][2] - a[0][2] * a[2][0];
    ia[1][2] = a[0][2] * a[1][0] - a[0][0] * a[1][2];
    ia[2][0] = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    ia[2][1] = a[0][1] * a[2][0] - a[0][0] * a[2][1];
    ia[2][2] = a[0][0] * a[1][1] - a[0][1] * a[1][0];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) ia[i][j] /= deta;
}

/* Gamma (src/init.c:541-591) applied as src/evolution.c:1373-1387 */
static void green_point(const orc_job *J, int v, const t2 sr, const t2 si, t2 orr, t2 oi)
{
    double G[3][3][3][3];
    const double *n = J->g[v];
    if (v == 0) {
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                for (int c = 0; c < 3; c++)
                    for (int d = 0; d < 3; d++)
                        G[a][b][c][d] = (J->ElastBC == 0) ? -1.0 * J->S0[a][b][c][d] : 0.0;
    } else {
        double io[3][3], om[3][3];
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) {
                double s = 0.0;
                for (int j = 0; j < 3; j++)
                    for (int l = 0; l < 3; l++) s += J->C0[i][j][k][l] * n[j] * n[l];
                io[i][k] = s;
            }
        m3inv(io, om);
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                for (int c = 0; c < 3; c++)
                    for (int d = 0; d < 3; d++) G[a][b][c][d] = -1.0 * om[a][c] * n[b] * n[d];
    }
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double r = 0.0, im = 0.0;
            for (int c = 0; c < 3; c++)
                for (int d = 0; d < 3; d++) {
                    r += G[a][b][c][d] * sr[c][d];
                    im += G[a][b][c][d] * si[c][d];
                }
            orr[a][b] = r;
            oi[a][b] = im;
        }
}

void orc_job_green(const orc_job *J, const double *ksig_r, const double *ksig_i, double *out_r,
                   double *out_i)
{
#pragma omp parallel for schedule(static)
    for (int v = 0; v < J->Nxyz; v++)
        green_point(J, v, (const double(*)[3])(ksig_r + 9 * (size_t)v),
                    (const double(*)[3])(ksig_i + 9 * (size_t)v), (double(*)[3])(out_r + 9 * (size_t)v),
                    (double(*)[3])(out_i + 9 * (size_t)v));
}

/* get_smacro (src/evolution.c:473-547), Q3 fixed, from the column sums of
 * the stress over all voxels (sum) and over phase-1 voxels (sum1) */
void orc_smacro_core(const double sum[3][3], const double sum1[3][3], double WGT, double Wgt_ph1,
                     const double S0[3][3][3][3], const double Scauchy[3][3], const int VelGrad_BC_Flag[6],
                     const int Stress_BC_Flag[6], double SigAvg[3][3], double SigAvg1[3][3],
                     double dDisGradAvg[3][3], double dDisGradAvg_acum[3][3], double *s_vm_out,
                     double *s_vm1_out)
{
    static const int IJV[2][6] = {{0, 1, 2, 1, 0, 0}, {0, 1, 2, 2, 2, 1}};
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            SigAvg[a][b] = sum[a][b] * WGT;
            SigAvg1[a][b] = sum1[a][b] * (WGT / Wgt_ph1);
        }
    for (int i = 0; i < 6; i++) {
        int ii = IJV[0][i], jj = IJV[1][i];
        dDisGradAvg[ii][jj] = 0.0;
        if (VelGrad_BC_Flag[i] == 0)
            for (int k = 0; k < 6; k++) {
                int kk = IJV[0][k], ll = IJV[1][k];
                dDisGradAvg[ii][jj] += S0[ii][jj][kk][ll] * Stress_BC_Flag[k] * (Scauchy[kk][ll] - SigAvg[kk][ll]);
            }
    }
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) dDisGradAvg_acum[a][b] += dDisGradAvg[a][b];
    double s6[6], s5[5], dev[3][3];
    orc_t2_to_v6((const double(*)[3])SigAvg, s6);
    for (int i = 0; i < 5; i++) s5[i] = s6[i];
    orc_v5_to_t2(s5, dev);
    double tr = dev[0][0] + dev[1][1] + dev[2][2];
    double s_vm = 0.0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            if (a != b) s_vm += dev[a][b] * dev[a][b];
            else s_vm += (dev[a][b] - tr / 3.0) * (dev[a][b] - tr / 3.0);
        }
    *s_vm_out = sqrt(3. / 2. * s_vm);
    orc_t2_to_v6((const double(*)[3])SigAvg1, s6);
    for (int i = 0; i < 5; i++) s5[i] = s6[i];
    orc_v5_to_t2(s5, dev);
    double s_vm1 = 0.0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) s_vm1 += dev[a][b] * dev[a][b];
    *s_vm1_out = sqrt(3. / 2. * s_vm1);
}

static void get_smacro(orc_job *J)
{
    double sum[3][3], sum1[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double s = 0.0, s1 = 0.0;
            LOOPV {
                s += J->Sig[v][a][b];
                if (J->phase[v] == 1) s1 += J->Sig[v][a][b];
            }
            sum[a][b] = s;
            sum1[a][b] = s1;
        }
    orc_smacro_core((const double(*)[3])sum, (const double(*)[3])sum1, J->WGT, J->Wgt_ph1,
                    (const double(*)[3][3][3])J->S0, (const double(*)[3])J->Scauchy, J->VelGrad_BC_Flag,
                    J->Stress_BC_Flag, J->SigAvg, J->SigAvg1, J->dDisGradAvg, J->dDisGradAvg_acum, &J->s_vm,
                    &J->s_vm1);
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* one Green iteration (src/evolution.c:1355-1462) */
/* the rate-law context of crystal voxel v: phase constants, the DRX term of
 * the passing stress (src/constitutive.c:584-586; t_last stays at its
 * initial value before any DRX event) and the voxel's densities           */
static void dd_context(orc_job *J, int v, orc_dd_phase *ph, nrctx *dd)
{
    const int p = J->phase[v] - 1;
    const double kT = orc_dd_kT(J->T_K);
    ph->c1 = J->Ck[0][p];
    ph->c2 = J->Ck[1][p];
    ph->c3 = J->Ck[2][p];
    ph->G = J->Shear_G[p];
    ph->q_slip = J->Q_slip[p] * MY_J_IN_EV_REF / kT;
    ph->kT = kT;
    ph->drx = J->DeltaTau_DRX * exp((J->t_last[v] - J->TimeTot) / J->TimeStep / J->tau_DRX);
    ph->bbB = J->bb_B[p];
    ph->chi = (const double(*)[ORC_NSYSMX])J->Chi[p];
    ph->cth = (const double(*)[ORC_NSYSMX][ORC_NSYSMX])J->cth[p];
    memset(dd, 0, sizeof(*dd));
    dd->dd = ph;
    dd->rho_s = J->rho_s[v];
    dd->rho_g1 = J->rho_g1[v];
    dd->rho_g2 = J->rho_g2[v];
    dd->rho_g3 = J->rho_g3[v];
    dd->rho_m = J->rho_m[v];
    dd->rho_F = J->rho_F[v];
    dd->rho_P = J->rho_P[v];
}

/* Gradient_ExchangeGrainID, one slab (src/constitutive.c:724-873): whether
 * the +x, -x, +y, -y, +z neighbours share the voxel's grain; the sixth flag
 * compares with the -y neighbour, as the reference does (Q20)             */
static void dd_grain_checks(orc_job *J)
{
    const int nx = J->CellDim[0], ny = J->CellDim[1], nz = J->CellDim[2];
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++)
            for (int z = 0; z < nz; z++) {
                int xp = (x + 1) % nx, xm = (x + nx - 1) % nx, yp = (y + 1) % ny,
                    ym = (y + ny - 1) % ny, zp = (z + 1) % nz;
                int v = (x * ny + y) * nz + z, g = J->grain[v];
                J->gbc[v][0] = g == J->grain[(xp * ny + y) * nz + z];
                J->gbc[v][1] = g == J->grain[(xm * ny + y) * nz + z];
                J->gbc[v][2] = g == J->grain[(x * ny + yp) * nz + z];
                J->gbc[v][3] = g == J->grain[(x * ny + ym) * nz + z];
                J->gbc[v][4] = g == J->grain[(x * ny + y) * nz + zp];
                J->gbc[v][5] = g == J->grain[(x * ny + ym) * nz + z];
            }
}

/* one-sided or central difference of f along one axis, by the two grain
 * checks (src/constitutive.c:1012-1053): 0 when both neighbours belong to
 * other grains                                                             */
static double dd_diff(int same_p, int same_m, double f, double fp, double fm)
{
    if (same_p == 0 && same_m == 1) return (f - fm);
    if (same_p == 1 && same_m == 0) return (fp - f);
    if (same_p == 1 && same_m == 1) return (fp - fm) / 2.0;
    return 0.0;
}

/* Gradient_ShearRate, one slab (src/constitutive.c:875-1108): the serial
 * exchange buffers hold the opposite x faces, so every voxel sees periodic
 * neighbours                                                               */
static void dd_gradient(orc_job *J)
{
    const int nx = J->CellDim[0], ny = J->CellDim[1], nz = J->CellDim[2];
#pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++)
            for (int z = 0; z < nz; z++) {
                int xp = (x + 1) % nx, xm = (x + nx - 1) % nx, yp = (y + 1) % ny,
                    ym = (y + ny - 1) % ny, zp = (z + 1) % nz, zm = (z + nz - 1) % nz;
                int v = (x * ny + y) * nz + z;
                int vxp = (xp * ny + y) * nz + z, vxm = (xm * ny + y) * nz + z;
                int vyp = (x * ny + yp) * nz + z, vym = (x * ny + ym) * nz + z;
                int vzp = (x * ny + y) * nz + zp, vzm = (x * ny + y) * nz + zm;
                const unsigned char *k = J->gbc[v];
                for (int i = 0; i < ORC_NSYSMX; i++) {
                    double f = J->gamdot[v][i];
                    J->gadot_grad[v][i][0] = dd_diff(k[0], k[1], f, J->gamdot[vxp][i], J->gamdot[vxm][i]);
                    J->gadot_grad[v][i][1] = dd_diff(k[2], k[3], f, J->gamdot[vyp][i], J->gamdot[vym][i]);
                    J->gadot_grad[v][i][2] = dd_diff(k[4], k[5], f, J->gamdot[vzp][i], J->gamdot[vzm][i]);
                }
            }
}

/* DislocationEvolution (src/constitutive.c:1346-1520): SSD production,
 * annihilation and climb, the diagnostic dipole / forest sums, first-order
 * GND rates from the shear-rate gradient.  The climb branch for
 * istep > 20000000 is not restated (Q21).                                  */
static void dd_evolution(orc_job *J)
{
    const double kT = orc_dd_kT(J->T_K), dt = J->TimeStep;
#pragma omp parallel for schedule(static)
    for (int v = 0; v < J->Nxyz; v++) {
        const int jph = J->phase[v];
        if (J->Type_phases[jph - 1]) continue;
        const int p = jph - 1;
        const double c_4 = J->Ck[3][p], c_5 = J->Ck[4][p], c_6 = J->Ck[5][p], c_7 = J->Ck[6][p];
        const double q_bulk = J->Q_bulk[p] * MY_J_IN_EV_REF / kT;
        const double shear_G = J->Shear_G[p], Poisson_nu = J->ElastConst[0][3];
        double sig6[6], rss[ORC_NSYSMX], hydro[ORC_NSYSMX];
        orc_t2_to_v6(J->Sig[v], sig6);
        const int ns = J->nSYS[p];
        for (int i = 0; i < ns; i++) {
            const double *sc = J->Schm_gr[v][i];
            rss[i] = sc[0] * sig6[0] + sc[1] * sig6[1] + sc[2] * sig6[2] + sc[3] * sig6[3] +
                     sc[4] * sig6[4];
            hydro[i] = (sig6[0] + sig6[1] + sig6[2]) / 3.0;
        }
        for (int i = 0; i < ns; i++) {
            const double b = J->bb_B[p][i];
            double d_dipole = sqrt(3.0) * shear_G * b / 16.0 / PI_REF / (1 - Poisson_nu) /
                              (fabs(rss[i]) + EPSILON_DD);
            double dot_SSD = 0.0;
            if (J->istep <= 2000000) {
                double rs = J->rho_s[v][i];
                J->rho_dipole[v][i] += fabs(J->gamdot[v][i]) * (c_6 * (d_dipole)*J->rho_m[v][i]) * dt;
                J->rho_forest[v][i] += fabs(J->gamdot[v][i]) * (c_4 * sqrt(J->rho_F[v][i])) * dt;
                dot_SSD = fabs(J->gamdot[v][i]) *
                              (c_4 * sqrt(J->rho_F[v][i]) + c_6 * (d_dipole)*J->rho_m[v][i] - c_5 * rs) -
                          c_7 * 1.6 * 1E12 * 1E-9 * rs * rs -
                          4 * 1E12 * (d_dipole) * (J->D_0 / b) * pow(rs, 2.0) *
                              exp((fabs(hydro[i]) * b * b * b * 100 / (1.38 * J->T_K)) - 1.0) *
                              exp(-1.0 * q_bulk);
                J->rho_s[v][i] += dot_SSD * dt;
            }
            if (J->istep <= 200000000) {
                const double *g = J->gadot_grad[v][i];
                const double *bp = J->bp[p][i], *bb = J->bb[p][i];
                double gnd1 = (g[0] * bp[0] + g[1] * bp[1] + g[2] * bp[2]);
                double dot_GND = gnd1 / b / J->L0;
                J->rho_g1[v][i] += dot_GND * dt;
                gnd1 = (g[0] * bb[0] + g[1] * bb[1] + g[2] * bb[2]);
                dot_GND = -1.0 * gnd1 / b / J->L0;
                J->rho_g2[v][i] += dot_GND * dt;
                gnd1 = 0.0;
                dot_GND = gnd1;
                J->rho_g3[v][i] += dot_GND * dt;
            }
        }
    }
}

/* AvgDislocation (src/constitutive.c:1669-1780): sums over crystal voxels
 * and systems, divided by the voxel count (gas voxels count as zero)     */
static void dd_averages(orc_job *J)
{
    double a[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 7; k++)
        for (int v = 0; v < J->Nxyz; v++) {
            const int jph = J->phase[v];
            if (J->Type_phases[jph - 1]) continue;
            for (int i = 0; i < J->nSYS[jph - 1]; i++) {
                double x;
                switch (k) {
                case 0: x = J->rho_F[v][i]; break;
                case 1: x = J->rho_P[v][i]; break;
                case 2: x = J->rho_s[v][i]; break;
                case 3: x = J->rho_m[v][i]; break;
                case 4:
                    x = sqrt(J->rho_g1[v][i] * J->rho_g1[v][i] + J->rho_g2[v][i] * J->rho_g2[v][i] +
                             J->rho_g3[v][i] * J->rho_g3[v][i]);
                    break;
                case 5: x = J->rho_dipole[v][i]; break;
                default: x = J->rho_forest[v][i]; break;
                }
                a[k] += x;
            }
        }
    for (int k = 0; k < 7; k++) J->dd_avg[k] = a[k] * (1.0 / J->Nxyz);
}

static int green_iteration(orc_job *J)
{
    const int N = J->Nxyz;
    const int nx = J->CellDim[0], ny = J->CellDim[1], nz = J->CellDim[2];
    /* forward transforms: Sig is exactly symmetric, so FFT(Sig_ab) == FFT(Sig_ba) */
    for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++) {
            for (int v = 0; v < N; v++) {
                J->kbuf[2 * v] = J->Sig[v][a][b];
                J->kbuf[2 * v + 1] = 0.0;
            }
            orc_fft3d(J->kbuf, nx, ny, nz, -1);
            for (int v = 0; v < N; v++) {
                J->ksr[v][a][b] = J->ksr[v][b][a] = J->kbuf[2 * v];
                J->ksi[v][a][b] = J->ksi[v][b][a] = J->kbuf[2 * v + 1];
            }
        }
#pragma omp parallel for schedule(static)
    for (int v = 0; v < N; v++) {
        t2 orr, oi;
        green_point(J, v, J->ksr[v], J->ksi[v], orr, oi);
        memcpy(J->ksr[v], orr, sizeof(t2));
        memcpy(J->ksi[v], oi, sizeof(t2));
    }
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            for (int v = 0; v < N; v++) {
                J->kbuf[2 * v] = J->ksr[v][a][b];
                J->kbuf[2 * v + 1] = J->ksi[v][a][b];
            }
            orc_fft3d(J->kbuf, nx, ny, nz, +1);
            double dd = J->dDisGradAvg[a][b];
            for (int v = 0; v < N; v++)
                J->DisGrad[v][a][b] += dd + J->kbuf[2 * v] / N;
        }
    /* UpdateStress over voxels; error sums taken in voxel order afterwards */
    int st_all = ORC_OK;
#pragma omp parallel for schedule(dynamic, 64)
    for (int v = 0; v < N; v++) {
        int jp = J->phase[v];
        J->nr_v[v] = 0;
        J->err_s_v[v] = J->err_e_v[v] = 0.0;
        if (!J->Type_phases[jp - 1]) {
            orc_dd_phase ph;
            nrctx dd;
            if (J->law == ORC_LAW_DD) dd_context(J, v, &ph, &dd);
            int st = update_stress(
                (const double(*)[6])J->sg[v], (const double(*)[6])J->C066, J->TimeStep, J->Sig[v],
                (const double(*)[3])J->Eps[v], (const double(*)[3])J->DisGrad[v], J->nSYS[jp - 1],
                (const double(*)[5])J->Schm_gr[v], (const double(*)[2])J->crss[v], J->nSRS[jp - 1],
                J->gam0[jp - 1], J->Edot[v], J->gamdot[v], &J->err_s_v[v], &J->err_e_v[v],
                &J->nr_v[v], J->law == ORC_LAW_DD ? &dd : NULL);
            if (st) {
#pragma omp critical
                st_all = st;
            }
        } else {
            memset(J->Sig[v], 0, sizeof(t2));
            memset(J->Edot[v], 0, sizeof(t2));
        }
    }
    if (st_all) return st_all;
    double es = 0.0, ee = 0.0;
    for (int v = 0; v < N; v++) {
        if (!J->Type_phases[J->phase[v] - 1]) {
            es += J->err_s_v[v];
            ee += J->err_e_v[v];
            J->newton_iters += J->nr_v[v];
        }
    }
    /* a non-finite stress or rate makes these sums non-finite: stop, as the
     * reference's isnan asserts do (src/constitutive.c:617,653,700) */
    if (!isfinite(es) || !isfinite(ee)) return ORC_ERR_NAN;
    J->Err_e = ee * J->WGT;
    J->Err_s = es * J->WGT;
    get_smacro(J);
    if (fabs(J->e_vm) < 1E-5) J->Err_e = 0.0;
    else J->Err_e /= J->e_vm;
    J->Err_s /= J->s_vm;
    return ORC_OK;
}

/* one increment: src/evolution.c:1318-1547 + reconstructed harden (Q1) */
static int increment(orc_job *J, int forced)
{
    const int N = J->Nxyz;
    const double dt = J->TimeStep;
    if (J->CREEP_FLAG == 1 && J->e_vm > 0.3) return -1;
    for (int v = 0; v < N; v++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                J->VelGrad[v][a][b] = J->DisGrad[v][a][b];
                if (J->CREEP_FLAG == 1) J->DisGrad[v][a][b] = J->DisGradAvg_actual[a][b];
                else J->DisGrad[v][a][b] += J->Udot[a][b] * dt;
            }
    if (J->istep == 1 || J->Update_Flag == 1) {
        for (int v = 0; v < N; v++) {
            int jp = J->phase[v];
            if (!J->Type_phases[jp - 1])
                orc_rotate_schmid(J->TranMat[v], J->nSYS[jp - 1], J->Schm_xt[jp - 1], J->Schm_gr[v]);
        }
    }
    memset(J->dDisGradAvg, 0, sizeof(t2));
    memset(J->dDisGradAvg_acum, 0, sizeof(t2));
    int iter = 0;
    J->Err_e = 2. * J->Err;
    J->Err_s = 2. * J->Err;
    double t0 = now_s();
    while (forced > 0 ? (iter < forced)
                      : ((iter < J->IterMax) && (MAXD(fabs(J->Err_s), fabs(J->Err_e)) > J->Err))) {
        iter++;
        int st = green_iteration(J);
        if (st) return -st;
    }
    J->green_time += now_s() - t0;
    J->green_iters += iter;
    /* epilogue */
    for (int v = 0; v < N; v++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                J->VelGrad[v][a][b] = (J->DisGrad[v][a][b] - J->VelGrad[v][a][b]) / dt;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            J->DisGradAvg_actual[a][b] = J->DisGradAvg[a][b] + J->dDisGradAvg_acum[a][b];
            J->VelGradAvg[a][b] = (J->DisGradAvg_actual[a][b] - J->DisGradAvg_t[a][b]) / dt;
        }
    J->e_vm = orc_vonmises(J->DisGradAvg_actual);
    J->d_vm = orc_vonmises(J->VelGradAvg);
    J->TimeTot += dt;
    memcpy(J->DisGradAvg_t, J->DisGradAvg_actual, sizeof(t2));
    for (int v = 0; v < N; v++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) J->Eps[v][a][b] += J->Edot[v][a][b] * dt;
    t2 ea = {{0}}, da = {{0}};
    for (int v = 0; v < N; v++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                ea[a][b] += J->Eps[v][a][b] * J->WGT;
                da[a][b] += J->Edot[v][a][b] * J->WGT;
            }
    memcpy(J->EpsAvg, ea, sizeof(t2));
    memcpy(J->EdotAvg, da, sizeof(t2));
    J->evmp = 0.0;
    J->dvmp = 0.0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            J->evmp += J->EpsAvg[a][b] * J->EpsAvg[a][b];
            J->dvmp += J->EdotAvg[a][b] * J->EdotAvg[a][b];
        }
    J->evmp = sqrt(2. / 3. * J->evmp);
    J->dvmp = sqrt(2. / 3. * J->dvmp);
    if (J->Update_Flag == 1 && J->istep > 1) {
        /* update_orient (src/constitutive.c:196-273) */
#pragma omp parallel for schedule(static)
        for (int v = 0; v < N; v++) {
            int jp = J->phase[v];
            if (J->Type_phases[jp - 1]) continue;
            t2 rl, lp = {{0}}, rs, rot, tm;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    rl[a][b] = (J->VelGrad[v][a][b] - J->VelGrad[v][b][a]) / 2.;
                    tm[a][b] = J->TranMat[v][a][b];
                }
            for (int s = 0; s < J->nSYS[jp - 1]; s++) {
                double bs[3], ns[3];
                for (int i = 0; i < 3; i++) {
                    bs[i] = 0.0;
                    ns[i] = 0.0;
                    for (int j = 0; j < 3; j++) {
                        bs[i] += tm[i][j] * J->bb[jp - 1][s][j];
                        ns[i] += tm[i][j] * J->bn[jp - 1][s][j];
                    }
                }
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) lp[a][b] += bs[a] * ns[b] * J->gamdot[v][s];
            }
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) rs[a][b] = (lp[a][b] - lp[b][a]) / 2.;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++)
                    rot[a][b] = (J->Udot_a[a][b] + rl[a][b] - rs[a][b]) * dt;
            orc_orientation(tm, rot);
            memcpy(J->TranMat[v], tm, sizeof(t2));
        }
    }
    if (J->law == ORC_LAW_DD) {
        /* src/evolution.c:1551-1564: harden() is not called; the gradient
         * of the shear rates (Hard_Flag, istep > 1), then the densities */
        if (J->Hard_Flag == 1 && J->istep > 1 && J->istep <= 20000000) dd_gradient(J);
        dd_evolution(J);
        dd_averages(J);
    } else if (J->Hard_Flag == 1 && J->istep > 1) {
#pragma omp parallel for schedule(static)
        for (int v = 0; v < N; v++) {
            int jp = J->phase[v];
            if (J->Type_phases[jp - 1]) continue;
            orc_harden_voxel(J->nSYS[jp - 1], J->gamdot[v], dt,
                             (const double(*)[ORC_NSYSMX])J->Hard[jp - 1],
                             (const double(*)[3])J->tau[jp - 1], (const double(*)[2])J->thet[jp - 1],
                             J->crss[v], &J->gamacum[v]);
        }
    }
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) J->DisGradAvg[a][b] = J->DisGradAvg_t[a][b] + J->Udot[a][b] * dt;
    J->istep++;
    return iter;
}

int orc_job_run(orc_job *J, int n, int *iters, const int *forced)
{
    for (int k = 0; k < n; k++) {
        int it = increment(J, forced ? forced[k] : 0);
        if (it < 0) {
            J->status = -it;
            return (it == -1) ? ORC_OK : -it;
        }
        if (iters) iters[k] = it;
    }
    return ORC_OK;
}

int orc_job_nvox(const orc_job *J) { return J->Nxyz; }

int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
/* threads of the OpenMP loops over voxels (results do not depend on it) */
void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
void orc_job_dims(const orc_job *J, int d[3]) { memcpy(d, J->CellDim, sizeof(J->CellDim)); }
int *orc_job_grain(orc_job *J) { return J->grain; }
int *orc_job_phase(orc_job *J) { return J->phase; }
int *orc_job_newton_counts(orc_job *J) { return J->nr_v; }
double orc_job_newton_iters(const orc_job *J) { return J->newton_iters; }

double *orc_job_field(orc_job *J, int which)
{
    switch (which) {
    case 0: return &J->Sig[0][0][0];
    case 1: return &J->Eps[0][0][0];
    case 2: return &J->Edot[0][0][0];
    case 3: return &J->DisGrad[0][0][0];
    case 4: return &J->VelGrad[0][0][0];
    case 5: return &J->TranMat[0][0][0];
    case 6: return &J->crss[0][0][0];
    case 7: return &J->gamdot[0][0];
    case 8: return J->gamacum;
    case 9: return J->rho_s ? &J->rho_s[0][0] : NULL;
    case 10: return J->rho_m ? &J->rho_m[0][0] : NULL;
    case 11: return J->rho_g1 ? &J->rho_g1[0][0] : NULL;
    case 12: return J->rho_g2 ? &J->rho_g2[0][0] : NULL;
    case 13: return J->rho_g3 ? &J->rho_g3[0][0] : NULL;
    case 14: return J->rho_F ? &J->rho_F[0][0] : NULL;
    case 15: return J->rho_P ? &J->rho_P[0][0] : NULL;
    case 16: return J->gadot_grad ? &J->gadot_grad[0][0][0] : NULL;
    case 17: return J->rho_dipole ? &J->rho_dipole[0][0] : NULL;
    case 18: return J->rho_forest ? &J->rho_forest[0][0] : NULL;
    }
    return NULL;
}

int orc_job_dd_evolve(orc_job *J)
{
    if (J->law != ORC_LAW_DD) return -1;
    dd_evolution(J);
    return ORC_OK;
}

int orc_job_dd_gradient(orc_job *J)
{
    if (J->law != ORC_LAW_DD) return -1;
    dd_gradient(J);
    return ORC_OK;
}

int orc_job_dd_tables(const orc_job *J, double *shear_G, double *bbB, double *bp, double *chi,
                      double *cth)
{
    memcpy(shear_G, J->Shear_G, sizeof(J->Shear_G));
    memcpy(bbB, J->bb_B, sizeof(J->bb_B));
    memcpy(bp, J->bp, sizeof(J->bp));
    memcpy(chi, J->Chi, sizeof(J->Chi));
    memcpy(cth, J->cth, sizeof(J->cth));
    return J->law;
}

/* layout of the macro vector, shared with oracle/oracle.py */
void orc_job_macro(const orc_job *J, double *o)
{
    memset(o, 0, 72 * sizeof(double));
    o[0] = J->TimeTot; o[1] = J->e_vm; o[2] = J->evmp; o[3] = J->d_vm; o[4] = J->dvmp;
    o[5] = J->s_vm; o[6] = J->s_vm1; o[7] = J->Err_e; o[8] = J->Err_s; o[9] = J->Wgt_ph1;
    memcpy(o + 10, J->SigAvg, 9 * sizeof(double));
    memcpy(o + 19, J->SigAvg1, 9 * sizeof(double));
    memcpy(o + 28, J->DisGradAvg_t, 9 * sizeof(double));
    memcpy(o + 37, J->EpsAvg, 9 * sizeof(double));
    memcpy(o + 46, J->EdotAvg, 9 * sizeof(double));
    o[55] = J->green_time; o[56] = J->green_iters; o[57] = J->newton_iters;
    o[58] = J->istep; o[59] = J->status;
    memcpy(o + 60, J->dd_avg, 7 * sizeof(double));
}

void orc_job_setup_info(const orc_job *J, double c066[36], double s066[36], double c0[81],
                        double s0[81])
{
    memcpy(c066, J->C066, sizeof(J->C066));
    memcpy(s066, J->S066, sizeof(J->S066));
    memcpy(c0, J->C0, sizeof(J->C0));
    memcpy(s0, J->S0, sizeof(J->S0));
}

int orc_job_cgr(const orc_job *J, double *cgr)
{
    memcpy(cgr, J->C_gr, sizeof(double) * 36 * (size_t)J->Nxyz);
    return J->Nxyz;
}
