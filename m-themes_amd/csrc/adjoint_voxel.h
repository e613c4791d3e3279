// adjoint_voxel.h -- per-voxel pieces of the reverse pass, written once for
// the device kernels (adjoint.hip) and for the host checks that compare each
// of them with finite differences of its forward counterpart.
//
// Every function here restates one forward step of the loading increment and
// its transposed linearisation:
//   schmid    P_s = sym((T n_s) x (T b_s)) in the 5-D basis
//             (update_schmid, src/kinematics.c:430-463)
//   rates     gamdot_s = g0 |x|^n sgn x, x = P_s:sigma / crss_s^(+/-),
//             edot = sum P_s gamdot_s (StrainRate_eval, src/constitutive.c:5-67;
//             |x|^n sgn x in place of pow(x, n) for non-integer n)
//   voce      the analytic Voce increment of harden (src/constitutive.c:275-326)
//   orient    update_orient + Orientation (src/constitutive.c:142-273)
// The *_bar functions take the adjoint of the outputs and ADD the adjoint of
// the inputs into the arrays they are given.
#pragma once

#include <math.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define AHD __host__ __device__ __forceinline__
#else
#define AHD static inline
#endif

#define ADJ_NS 12

namespace adj {

constexpr double R2 = 0.70710678118654744, R3 = 0.57735026918962584, R6 = 0.40824829046386304;

// symmetric (11 22 33 23 13 12) -> B-basis coefficients (chg_basis opt 2)
AHD void sym_to_b(const double t[6], double v[6])
{
    v[0] = -R2 * t[0] + R2 * t[1];
    v[1] = -R6 * t[0] - R6 * t[1] + 2. * R6 * t[2];
    v[2] = 2. * R2 * t[3];
    v[3] = 2. * R2 * t[4];
    v[4] = 2. * R2 * t[5];
    v[5] = R3 * (t[0] + t[1] + t[2]);
}

// B-basis -> symmetric (chg_basis opt 1)
AHD void b_to_sym(const double v[6], double t[6])
{
    t[0] = -R2 * v[0] - R6 * v[1] + R3 * v[5];
    t[1] = R2 * v[0] - R6 * v[1] + R3 * v[5];
    t[2] = 2. * R6 * v[1] + R3 * v[5];
    t[3] = R2 * v[2];
    t[4] = R2 * v[3];
    t[5] = R2 * v[4];
}

// transpose of b_to_sym: adjoint of the symmetric components -> B adjoint
AHD void b_to_sym_T(const double tb[6], double vb[6])
{
    vb[0] = -R2 * tb[0] + R2 * tb[1];
    vb[1] = -R6 * tb[0] - R6 * tb[1] + 2. * R6 * tb[2];
    vb[2] = R2 * tb[3];
    vb[3] = R2 * tb[4];
    vb[4] = R2 * tb[5];
    vb[5] = R3 * (tb[0] + tb[1] + tb[2]);
}

AHD void schmid(const double T[3][3], const double bn[3], const double bb[3], double c[5])
{
    double tn[3], tb[3];
    for (int a = 0; a < 3; a++) {
        tn[a] = T[a][0] * bn[0] + T[a][1] * bn[1] + T[a][2] * bn[2];
        tb[a] = T[a][0] * bb[0] + T[a][1] * bb[1] + T[a][2] * bb[2];
    }
    double P00 = tn[0] * tb[0], P11 = tn[1] * tb[1], P22 = tn[2] * tb[2];
    double P12 = (tn[1] * tb[2] + tn[2] * tb[1]) / 2.0;
    double P02 = (tn[0] * tb[2] + tn[2] * tb[0]) / 2.0;
    double P01 = (tn[0] * tb[1] + tn[1] * tb[0]) / 2.0;
    c[0] = -R2 * P00 + R2 * P11;
    c[1] = -R6 * P00 - R6 * P11 + 2. * R6 * P22;
    c[2] = 2. * R2 * P12;
    c[3] = 2. * R2 * P02;
    c[4] = 2. * R2 * P01;
}

AHD void schmid_bar(const double T[3][3], const double bn[3], const double bb[3], const double cb[5],
                    double Tb[3][3])
{
    double tn[3], tb[3];
    for (int a = 0; a < 3; a++) {
        tn[a] = T[a][0] * bn[0] + T[a][1] * bn[1] + T[a][2] * bn[2];
        tb[a] = T[a][0] * bb[0] + T[a][1] * bb[1] + T[a][2] * bb[2];
    }
    double P00b = -R2 * cb[0] - R6 * cb[1], P11b = R2 * cb[0] - R6 * cb[1], P22b = 2. * R6 * cb[1];
    double P12b = 2. * R2 * cb[2], P02b = 2. * R2 * cb[3], P01b = 2. * R2 * cb[4];
    double tnb[3], tbb[3];
    tnb[0] = P00b * tb[0] + 0.5 * (P02b * tb[2] + P01b * tb[1]);
    tnb[1] = P11b * tb[1] + 0.5 * (P12b * tb[2] + P01b * tb[0]);
    tnb[2] = P22b * tb[2] + 0.5 * (P12b * tb[1] + P02b * tb[0]);
    tbb[0] = P00b * tn[0] + 0.5 * (P02b * tn[2] + P01b * tn[1]);
    tbb[1] = P11b * tn[1] + 0.5 * (P12b * tn[2] + P01b * tn[0]);
    tbb[2] = P22b * tn[2] + 0.5 * (P12b * tn[1] + P02b * tn[0]);
    for (int a = 0; a < 3; a++)
        for (int j = 0; j < 3; j++) Tb[a][j] += tnb[a] * bn[j] + tbb[a] * bb[j];
}

// Rate law of one voxel.  sig: B-basis stress (first 5 used); c[s]: Schmid
// vectors; crss[2s+d] (d = 0 forward, 1 backward).  Outputs gd[s], ed[6]
// (6th component 0) and, if dd != 0, the symmetric 6x6 d(edot)/d(sigma).
AHD void rates(int ns, const double c[ADJ_NS][5], const double *crss, const double *nrs,
               const double *g0, const double sig[6], double gd[ADJ_NS], double ed[6],
               double (*dd)[6])
{
    for (int i = 0; i < 6; i++) {
        ed[i] = 0.0;
        if (dd)
            for (int j = 0; j < 6; j++) dd[i][j] = 0.0;
    }
    for (int s = 0; s < ns; s++) {
        double r = 0.0;
        for (int j = 0; j < 5; j++) r += c[s][j] * sig[j];
        double it = 1.0 / (r < 0 ? crss[2 * s + 1] : crss[2 * s]);
        double x = r * it, a = fabs(x);
        double an1 = a > 0.0 ? pow(a, nrs[s] - 1.0) : 0.0;
        double g = g0[s] * an1 * a * (x > 0 ? 1.0 : -1.0);
        gd[s] = g;
        for (int i = 0; i < 5; i++) ed[i] += c[s][i] * g;
        if (dd) {
            double r1 = g0[s] * nrs[s] * an1 * it;
            for (int i = 0; i < 5; i++)
                for (int j = 0; j < 5; j++) dd[i][j] += c[s][i] * c[s][j] * r1;
        }
    }
}

// Adjoint of rates at fixed sigma for given gdb[s] (adjoint of gamdot) and
// edb[6] (adjoint of edot, B-basis).  Adds into sigb[6], cb[s][5], crssb[24]
// and nb[s] (d/d n_s).
AHD void rates_bar(int ns, const double c[ADJ_NS][5], const double *crss, const double *nrs,
                   const double *g0, const double sig[6], const double *gdb, const double edb[6],
                   double sigb[6], double cb[ADJ_NS][5], double *crssb, double *nb)
{
    for (int s = 0; s < ns; s++) {
        double r = 0.0;
        for (int j = 0; j < 5; j++) r += c[s][j] * sig[j];
        const int sel = r < 0 ? 2 * s + 1 : 2 * s;
        double it = 1.0 / crss[sel];
        double x = r * it, a = fabs(x);
        double an1 = a > 0.0 ? pow(a, nrs[s] - 1.0) : 0.0;
        double g = g0[s] * an1 * a * (x > 0 ? 1.0 : -1.0);
        double gt = gdb ? gdb[s] : 0.0;
        for (int i = 0; i < 5; i++) {
            gt += c[s][i] * edb[i];
            cb[s][i] += g * edb[i];
        }
        double xb = gt * g0[s] * nrs[s] * an1;
        if (a > 0.0) nb[s] += gt * g * log(a);
        double rb = xb * it;
        crssb[sel] += -xb * x * it;
        for (int i = 0; i < 5; i++) {
            sigb[i] += rb * c[s][i];
            cb[s][i] += rb * sig[i];
        }
    }
}

// Voce increment factor of harden (src/constitutive.c:297-316) for one
// system: G = accumulated shear, dg = shear increment of the step.
AHD double voce(double G, double dg, double tau0, double tau1, double th0, double th1)
{
    double tiny = 1.E-4 * tau0, v = 0.0;
    if (fabs(th0) > tiny) {
        v = th1 * dg;
        if (fabs(tau1) > tiny) {
            double f = fabs(th0 / tau1);
            double ei = exp(-1.0 * G * f), ed = exp(-1.0 * dg * f);
            v += -1. * (f * tau1 - th1) / f * ei * (ed - 1.0) -
                 th1 / f * ei * (ed * ((G + dg) * f + 1.0) - (G * f + 1.0));
        }
    }
    return v;
}

// Partial derivatives of voce: d[0] d/dG, d[1] d/ddg, d[2] d/dtau1,
// d[3] d/dth0, d[4] d/dth1 (the tau0 guard is a branch: no derivative).
AHD void voce_grad(double G, double dg, double tau0, double tau1, double th0, double th1, double d[5])
{
    for (int i = 0; i < 5; i++) d[i] = 0.0;
    double tiny = 1.E-4 * tau0;
    if (!(fabs(th0) > tiny)) return;
    d[1] = th1;
    d[4] = dg;
    if (!(fabs(tau1) > tiny)) return;
    double q = th0 / tau1, f = fabs(q), sq = q > 0 ? 1.0 : -1.0;
    double ei = exp(-G * f), ed = exp(-dg * f);
    // A = -(tau1 - th1/f) ei (ed - 1)
    double k = tau1 - th1 / f, em = ed - 1.0;
    double A_t1 = -ei * em;
    double A_th1 = ei * em / f;
    double A_f = -(th1 / (f * f)) * ei * em - k * (-G * ei * em + ei * (-dg * ed));
    double A_G = -k * (-f * ei) * em;
    double A_dg = -k * ei * (-f * ed);
    // B = -(th1/f) ei Q,  Q = ed ((G+dg) f + 1) - (G f + 1)
    double Q = ed * ((G + dg) * f + 1.0) - (G * f + 1.0);
    double Q_G = ed * f - f;
    double Q_dg = -f * ed * ((G + dg) * f + 1.0) + ed * f;
    double Q_f = -dg * ed * ((G + dg) * f + 1.0) + ed * (G + dg) - G;
    double B_th1 = -ei * Q / f;
    double B_f = th1 / (f * f) * ei * Q + (th1 / f) * G * ei * Q - (th1 / f) * ei * Q_f;
    double B_G = (th1 / f) * f * ei * Q - (th1 / f) * ei * Q_G;
    double B_dg = -(th1 / f) * ei * Q_dg;
    double df_dth0 = sq / tau1, df_dt1 = -sq * th0 / (tau1 * tau1);
    d[0] = A_G + B_G;
    d[1] += A_dg + B_dg;
    d[2] = A_t1 + (A_f + B_f) * df_dt1;
    d[3] = (A_f + B_f) * df_dth0;
    d[4] += A_th1 + B_th1;
}

struct HardParams {
    int ns;
    double dt;
    double hard[ADJ_NS][ADJ_NS];
    double tau[ADJ_NS][3];
    double thet[ADJ_NS][2];
};

// harden of one voxel: crss (24) and the accumulated shear G updated in place
AHD void harden(const HardParams &H, const double gd[ADJ_NS], double *crss, double &G)
{
    double dg = 0.0;
    for (int i = 0; i < H.ns; i++) dg += fabs(gd[i]) * H.dt;
    for (int i = 0; i < H.ns; i++) {
        double dtau = 0.0;
        for (int j = 0; j < H.ns; j++) dtau += H.hard[i][j] * fabs(gd[j]) * H.dt;
        double v = voce(G, dg, H.tau[i][0], H.tau[i][2], H.thet[i][0], H.thet[i][1]);
        if (dg != 0.0) {
            double inc = dtau * v / dg;
            crss[2 * i] += inc;
            crss[2 * i + 1] += inc;
        }
    }
    G += dg;
}

// Adjoint of harden.  crssb (24), Gb: adjoint of the outputs on entry; on
// return Gb holds the adjoint of the input G (crss passes through
// unchanged).  Adds into gdb[s] and pb[0..2] = d/d(tau1, theta0, theta1).
AHD void harden_bar(const HardParams &H, const double gd[ADJ_NS], double G, const double *crssb,
                    double &Gb, double *gdb, double pb[3])
{
    double dg = 0.0;
    for (int i = 0; i < H.ns; i++) dg += fabs(gd[i]) * H.dt;
    double dgb = Gb; // G_out = G + dg
    double Gin_b = Gb;
    double dtaub[ADJ_NS];
    for (int i = 0; i < H.ns; i++) dtaub[i] = 0.0;
    if (dg != 0.0) {
        for (int i = 0; i < H.ns; i++) {
            double dtau = 0.0;
            for (int j = 0; j < H.ns; j++) dtau += H.hard[i][j] * fabs(gd[j]) * H.dt;
            double v = voce(G, dg, H.tau[i][0], H.tau[i][2], H.thet[i][0], H.thet[i][1]);
            double gb = crssb[2 * i] + crssb[2 * i + 1];
            dtaub[i] = gb * v / dg;
            double vb = gb * dtau / dg;
            dgb += -gb * dtau * v / (dg * dg);
            double d[5];
            voce_grad(G, dg, H.tau[i][0], H.tau[i][2], H.thet[i][0], H.thet[i][1], d);
            Gin_b += vb * d[0];
            dgb += vb * d[1];
            pb[0] += vb * d[2];
            pb[1] += vb * d[3];
            pb[2] += vb * d[4];
        }
    }
    for (int j = 0; j < H.ns; j++) {
        double ab = dgb * H.dt;
        for (int i = 0; i < H.ns; i++) ab += dtaub[i] * H.hard[i][j] * H.dt;
        double sg = gd[j] > 0 ? 1.0 : (gd[j] < 0 ? -1.0 : 0.0);
        gdb[j] += ab * sg;
    }
    Gb = Gin_b;
}

struct OrientParams {
    int ns;
    double dt;
    double udot_a[9];
    double bn[ADJ_NS][3], bb[ADJ_NS][3];
};

// rotation increment omega (3x3) of update_orient from T, dW (23 13 12 of the
// antisymmetric displacement-gradient increment) and gamdot
AHD void orient_omega(const OrientParams &O, const double T[3][3], const double dW[3],
                      const double gd[ADJ_NS], double om[3][3])
{
    double lp[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int s = 0; s < O.ns; s++) {
        double tb[3], tn[3];
        for (int a = 0; a < 3; a++) {
            tb[a] = T[a][0] * O.bb[s][0] + T[a][1] * O.bb[s][1] + T[a][2] * O.bb[s][2];
            tn[a] = T[a][0] * O.bn[s][0] + T[a][1] * O.bn[s][1] + T[a][2] * O.bn[s][2];
        }
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) lp[a][b] += tb[a] * tn[b] * gd[s];
    }
    double w23 = dW[0] / O.dt, w13 = dW[1] / O.dt, w12 = dW[2] / O.dt;
    double rl[3][3] = {{0.0, w12, w13}, {-w12, 0.0, w23}, {-w13, -w23, 0.0}};
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            om[a][b] = (O.udot_a[a * 3 + b] + rl[a][b] - (lp[a][b] - lp[b][a]) / 2.) * O.dt;
}

// Orientation (src/constitutive.c:142-194): the rotation matrix of omega,
// with the reference's small-angle rule (|w| < 1e-6 divides by 1, not |w|)
AHD void orient_rot(const double om[3][3], double rot[3][3])
{
    double w[3] = {om[2][1], om[0][2], om[1][0]};
    double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    double wt = tan(wn / 2.0);
    if (fabs(wn) < 1E-6) wn = 1.0;
    for (int i = 0; i < 3; i++) w[i] *= wt / wn;
    double O[3][3] = {{0., -w[2], w[1]}, {w[2], 0., -w[0]}, {-w[1], w[0], 0.}};
    double q = wt * wt;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double o2 = O[a][0] * O[0][b] + O[a][1] * O[1][b] + O[a][2] * O[2][b];
            rot[a][b] = (double)(a == b) + 2. * (O[a][b] + o2) / (1. + q);
        }
}

AHD void orient(const OrientParams &O, double T[3][3], const double dW[3], const double gd[ADJ_NS])
{
    double om[3][3], rot[3][3], Tn[3][3];
    orient_omega(O, T, dW, gd, om);
    orient_rot(om, rot);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) Tn[a][b] = rot[a][0] * T[0][b] + rot[a][1] * T[1][b] + rot[a][2] * T[2][b];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) T[a][b] = Tn[a][b];
}

// Adjoint of orient.  Tb: adjoint of the output T on entry, of the input T on
// return.  Adds into gdb[s]; writes dWb[3].
AHD void orient_bar(const OrientParams &O, const double T[3][3], const double dW[3],
                    const double gd[ADJ_NS], double Tb[3][3], double *gdb, double dWb[3])
{
    double om[3][3], rot[3][3];
    orient_omega(O, T, dW, gd, om);
    orient_rot(om, rot);
    // T_out = rot T
    double rotb[3][3], Tin_b[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            rotb[a][b] = Tb[a][0] * T[b][0] + Tb[a][1] * T[b][1] + Tb[a][2] * T[b][2];
            Tin_b[a][b] = rot[0][a] * Tb[0][b] + rot[1][a] * Tb[1][b] + rot[2][a] * Tb[2][b];
        }
    // rot = I + 2 (Om + Om^2) / (1 + q)
    double w0[3] = {om[2][1], om[0][2], om[1][0]};
    double wn0 = sqrt(w0[0] * w0[0] + w0[1] * w0[1] + w0[2] * w0[2]);
    double wt = tan(wn0 / 2.0);
    const bool small = fabs(wn0) < 1E-6;
    double s = small ? wt : wt / wn0;
    double w[3] = {w0[0] * s, w0[1] * s, w0[2] * s};
    double Om[3][3] = {{0., -w[2], w[1]}, {w[2], 0., -w[0]}, {-w[1], w[0], 0.}};
    double q = wt * wt, den = 1. + q;
    double Omb[3][3], qb = 0.0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double o2 = Om[a][0] * Om[0][b] + Om[a][1] * Om[1][b] + Om[a][2] * Om[2][b];
            qb += -2. * rotb[a][b] * (Om[a][b] + o2) / (den * den);
            // d<Rb, Om Om>/dOm = Rb Om^T + Om^T Rb
            double t1 = rotb[a][0] * Om[b][0] + rotb[a][1] * Om[b][1] + rotb[a][2] * Om[b][2];
            double t2 = Om[0][a] * rotb[0][b] + Om[1][a] * rotb[1][b] + Om[2][a] * rotb[2][b];
            Omb[a][b] = 2. / den * (rotb[a][b] + t1 + t2);
        }
    double wb[3] = {Omb[2][1] - Omb[1][2], Omb[0][2] - Omb[2][0], Omb[1][0] - Omb[0][1]};
    double sb = wb[0] * w0[0] + wb[1] * w0[1] + wb[2] * w0[2];
    double w0b[3] = {wb[0] * s, wb[1] * s, wb[2] * s};
    double dwt = 0.5 * (1.0 + wt * wt);
    double ds = small ? dwt : (dwt * wn0 - wt) / (wn0 * wn0);
    double wnb = sb * ds + qb * 2.0 * wt * dwt;
    if (wn0 > 0.0)
        for (int i = 0; i < 3; i++) w0b[i] += wnb * w0[i] / wn0;
    double omb[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    omb[2][1] = w0b[0];
    omb[0][2] = w0b[1];
    omb[1][0] = w0b[2];
    // om = (udot_a + rl - skew(lp)) dt ; rl from dW / dt
    dWb[0] = -omb[2][1] + omb[1][2];
    dWb[1] = omb[0][2] - omb[2][0];
    dWb[2] = omb[0][1] - omb[1][0];
    double lpb[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) lpb[a][b] = -0.5 * O.dt * (omb[a][b] - omb[b][a]);
    for (int s2 = 0; s2 < O.ns; s2++) {
        double tb[3], tn[3];
        for (int a = 0; a < 3; a++) {
            tb[a] = T[a][0] * O.bb[s2][0] + T[a][1] * O.bb[s2][1] + T[a][2] * O.bb[s2][2];
            tn[a] = T[a][0] * O.bn[s2][0] + T[a][1] * O.bn[s2][1] + T[a][2] * O.bn[s2][2];
        }
        double g = 0.0, tbb[3] = {0, 0, 0}, tnb[3] = {0, 0, 0};
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                g += lpb[a][b] * tb[a] * tn[b];
                tbb[a] += gd[s2] * lpb[a][b] * tn[b];
                tnb[b] += gd[s2] * lpb[a][b] * tb[a];
            }
        gdb[s2] += g;
        for (int a = 0; a < 3; a++)
            for (int j = 0; j < 3; j++) Tin_b[a][j] += tbb[a] * O.bb[s2][j] + tnb[a] * O.bn[s2][j];
    }
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) Tb[a][b] = Tin_b[a][b];
}

// C_t = (S + dt D)^-1 for the symmetric 6x6 S (B-basis compliance) and D
// (d edot / d sigma); Cholesky of the symmetrised matrix.  Returns false if
// it is not positive definite.
AHD bool tangent(const double S[6][6], const double D[6][6], double dt, double Ct[6][6])
{
    double L[6][6];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) L[i][j] = 0.5 * (S[i][j] + S[j][i]) + dt * 0.5 * (D[i][j] + D[j][i]);
    for (int j = 0; j < 6; j++) {
        double d = L[j][j];
        for (int k = 0; k < j; k++) d -= L[j][k] * L[j][k];
        if (!(d > 0.0)) return false;
        d = sqrt(d);
        L[j][j] = d;
        for (int i = j + 1; i < 6; i++) {
            double s = L[i][j];
            for (int k = 0; k < j; k++) s -= L[i][k] * L[j][k];
            L[i][j] = s / d;
        }
    }
    // inverse of L, then Ct = L^-T L^-1
    double Li[6][6];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) Li[i][j] = 0.0;
    for (int j = 0; j < 6; j++) {
        Li[j][j] = 1.0 / L[j][j];
        for (int i = j + 1; i < 6; i++) {
            double s = 0.0;
            for (int k = j; k < i; k++) s -= L[i][k] * Li[k][j];
            Li[i][j] = s / L[i][i];
        }
    }
    for (int i = 0; i < 6; i++)
        for (int j = i; j < 6; j++) {
            double s = 0.0;
            for (int k = j; k < 6; k++) s += Li[k][i] * Li[k][j];
            Ct[i][j] = s;
            Ct[j][i] = s;
        }
    return true;
}

} // namespace adj
