// hostcheck.cpp -- host builds of device code that the tests check on the
// CPU, behind plain C entry points (test infrastructure, not part of the
// libraries):
//   * the per-voxel adjoint functions (m-themes_amd/csrc/adjoint_voxel.h,
//     the same source the device kernels compile), so that each transposed
//     linearisation can be checked against finite differences of its
//     forward function;
//   * the Newton kernel's line-search backtracking rule
//     (m-themes_amd/csrc/linesearch.h), checked against the oracle's and
//     against the models it minimises on every branch;
//   * the FP32 wave-vector directions and Nyquist split test
//     (m-themes_amd/csrc/spectral.h), checked against the FP64 ones over
//     whole half spectra;
//   * the Green loop control's get_smacro step
//     (m-themes_amd/csrc/loopctl_step.h), checked against the oracle's bit
//     for bit;
//   * the Green operator at one wave vector of the half spectrum
//     (m-themes_amd/csrc/green_point.h, FP64 and FP32), checked against the
//     oracle's full-grid operator.
#include <string.h>

#include "adjoint_voxel.h"
#include "green_point.h"
#include "linesearch.h"
#include "loopctl_step.h"
#include "spectral.h"

using namespace adj;

namespace {

void load_T(const double *t9, double T[3][3])
{
    for (int i = 0; i < 9; i++) T[i / 3][i % 3] = t9[i];
}

void store_T(const double T[3][3], double *t9)
{
    for (int i = 0; i < 9; i++) t9[i] = T[i / 3][i % 3];
}

HardParams hard_params(int ns, double dt, const double *hard, const double *tau, const double *thet)
{
    HardParams H;
    memset(&H, 0, sizeof(H));
    H.ns = ns;
    H.dt = dt;
    for (int i = 0; i < ns; i++) {
        for (int j = 0; j < ns; j++) H.hard[i][j] = hard[i * ns + j];
        for (int k = 0; k < 3; k++) H.tau[i][k] = tau[i * 3 + k];
        for (int k = 0; k < 2; k++) H.thet[i][k] = thet[i * 2 + k];
    }
    return H;
}

OrientParams orient_params(int ns, double dt, const double *udot_a, const double *bn, const double *bb)
{
    OrientParams O;
    memset(&O, 0, sizeof(O));
    O.ns = ns;
    O.dt = dt;
    for (int i = 0; i < 9; i++) O.udot_a[i] = udot_a[i];
    for (int s = 0; s < ns; s++)
        for (int k = 0; k < 3; k++) {
            O.bn[s][k] = bn[s * 3 + k];
            O.bb[s][k] = bb[s * 3 + k];
        }
    return O;
}

} // namespace

extern "C" {

void adjc_schmid(const double *t9, const double *bn, const double *bb, double *c5)
{
    double T[3][3];
    load_T(t9, T);
    schmid(T, bn, bb, c5);
}

void adjc_schmid_bar(const double *t9, const double *bn, const double *bb, const double *cb5, double *tb9)
{
    double T[3][3], Tb[3][3];
    load_T(t9, T);
    load_T(tb9, Tb);
    schmid_bar(T, bn, bb, cb5, Tb);
    store_T(Tb, tb9);
}

void adjc_rates(int ns, const double *c, const double *crss, const double *nrs, const double *g0,
                const double *sig, double *gd, double *ed, double *dd36)
{
    double C[ADJ_NS][5], D[6][6];
    for (int s = 0; s < ns; s++)
        for (int j = 0; j < 5; j++) C[s][j] = c[s * 5 + j];
    rates(ns, C, crss, nrs, g0, sig, gd, ed, dd36 ? D : nullptr);
    if (dd36)
        for (int i = 0; i < 36; i++) dd36[i] = D[i / 6][i % 6];
}

void adjc_rates_bar(int ns, const double *c, const double *crss, const double *nrs, const double *g0,
                    const double *sig, const double *gdb, const double *edb, double *sigb, double *cb,
                    double *crssb, double *nb)
{
    double C[ADJ_NS][5], CB[ADJ_NS][5];
    for (int s = 0; s < ns; s++)
        for (int j = 0; j < 5; j++) {
            C[s][j] = c[s * 5 + j];
            CB[s][j] = cb[s * 5 + j];
        }
    rates_bar(ns, C, crss, nrs, g0, sig, gdb, edb, sigb, CB, crssb, nb);
    for (int s = 0; s < ns; s++)
        for (int j = 0; j < 5; j++) cb[s * 5 + j] = CB[s][j];
}

double adjc_voce(double G, double dg, double tau0, double tau1, double th0, double th1)
{
    return voce(G, dg, tau0, tau1, th0, th1);
}

void adjc_voce_grad(double G, double dg, double tau0, double tau1, double th0, double th1, double *d5)
{
    voce_grad(G, dg, tau0, tau1, th0, th1, d5);
}

void adjc_harden(int ns, double dt, const double *hard, const double *tau, const double *thet,
                 const double *gd, double *crss, double *G)
{
    HardParams H = hard_params(ns, dt, hard, tau, thet);
    harden(H, gd, crss, *G);
}

void adjc_harden_bar(int ns, double dt, const double *hard, const double *tau, const double *thet,
                     const double *gd, double G, const double *crssb, double *Gb, double *gdb, double *pb3)
{
    HardParams H = hard_params(ns, dt, hard, tau, thet);
    harden_bar(H, gd, G, crssb, *Gb, gdb, pb3);
}

void adjc_orient(int ns, double dt, const double *udot_a, const double *bn, const double *bb, double *t9,
                 const double *dW, const double *gd)
{
    OrientParams O = orient_params(ns, dt, udot_a, bn, bb);
    double T[3][3];
    load_T(t9, T);
    orient(O, T, dW, gd);
    store_T(T, t9);
}

void adjc_orient_bar(int ns, double dt, const double *udot_a, const double *bn, const double *bb,
                     const double *t9, const double *dW, const double *gd, double *tb9, double *gdb,
                     double *dWb)
{
    OrientParams O = orient_params(ns, dt, udot_a, bn, bb);
    double T[3][3], Tb[3][3];
    load_T(t9, T);
    load_T(tb9, Tb);
    orient_bar(O, T, dW, gd, Tb, gdb, dWb);
    store_T(Tb, tb9);
}

int adjc_tangent(const double *S36, const double *D36, double dt, double *Ct36)
{
    double S[6][6], D[6][6], C[6][6];
    for (int i = 0; i < 36; i++) {
        S[i / 6][i % 6] = S36[i];
        D[i / 6][i % 6] = D36[i];
    }
    bool ok = tangent(S, D, dt, C);
    for (int i = 0; i < 36; i++) Ct36[i] = C[i / 6][i % 6];
    return ok ? 0 : 1;
}

} // extern "C"

extern "C" double lsc_backtrack(double alam, double alam2, double f_new, double f2, double f_old,
                                double slope)
{
    return ls_backtrack(alam, alam2, f_new, f2, f_old, slope);
}

// Every wave vector k != 0 of the nx x ny x nz half spectrum through both
// forms of kspace::mode_dirs.  out: [0] wave vectors, [1] split in FP64,
// [2] split decisions that differ, [3] max |n_f32 - n_f64| over n1 and, where
// split, n2 (each up to sign: the Green operator is even in n).
extern "C" void ksc_sweep(int nx, int ny, int nz, double *out)
{
    double tot = 0, splits = 0, bad = 0, dmax = 0;
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++)
            for (int kz = 0; kz <= nz / 2; kz++) {
                if (x + y + kz == 0) continue;
                double d1[3], d2[3];
                float f1[3], f2[3];
                const bool sd = kspace::mode_dirs(x, y, kz, nx, ny, nz, d1, d2);
                const bool sf = kspace::mode_dirs(x, y, kz, nx, ny, nz, f1, f2);
                tot += 1;
                splits += sd;
                bad += sd != sf;
                for (int v = 0; v < (sd ? 2 : 1); v++) {
                    const double *d = v ? d2 : d1;
                    const float *f = v ? f2 : f1;
                    double ep = 0, em = 0;
                    for (int m = 0; m < 3; m++) {
                        ep = fmax(ep, fabs(f[m] - d[m]));
                        em = fmax(em, fabs(f[m] + d[m]));
                    }
                    dmax = fmax(dmax, fmin(ep, em));
                }
            }
    out[0] = tot;
    out[1] = splits;
    out[2] = bad;
    out[3] = dmax;
}

// lctl::smacro on one set of reduced sums: red[3..8] the stress sums over
// all voxels, red[9..14] over phase-1 voxels (11 22 33 23 13 12).  dD and
// dD_acum (9 each, row-major) are updated in place; out receives sig_avg (9),
// sig_avg1 (9), s_vm and s_vm1.
extern "C" void lcc_smacro(const double *s0, const double *scauchy, double wgt_ph1, double scale,
                           const int *velgrad_bc, const int *stress_bc, const double *red, double *dD,
                           double *dD_acum, double *out)
{
    CtlConst K;
    LoopCtl L;
    memset(&K, 0, sizeof(K));
    memset(&L, 0, sizeof(L));
    memcpy(K.s0, s0, sizeof(K.s0));
    memcpy(K.scauchy, scauchy, sizeof(K.scauchy));
    K.wgt_ph1 = wgt_ph1;
    K.scale = scale;
    for (int i = 0; i < 6; i++) {
        K.velgrad_bc[i] = velgrad_bc[i];
        K.stress_bc[i] = stress_bc[i];
    }
    memcpy(L.dD, dD, sizeof(L.dD));
    memcpy(L.dD_acum, dD_acum, sizeof(L.dD_acum));
    lctl::smacro(K, L, red);
    memcpy(dD, L.dD, sizeof(L.dD));
    memcpy(dD_acum, L.dD_acum, sizeof(L.dD_acum));
    memcpy(out, L.sig_avg, sizeof(L.sig_avg));
    memcpy(out + 9, L.sig_avg1, sizeof(L.sig_avg1));
    out[18] = L.s_vm;
    out[19] = L.s_vm1;
}


namespace {
struct cplx64 {
    double x, y;
};
struct cplx32 {
    float x, y;
};

template <typename R, typename C>
int green_point_as(const double *ga36, int x, int y, int kz, int nx, int ny, int nz, const double *s12,
                   double *e12, double *w6)
{
    R ga[36];
    for (int i = 0; i < 36; i++) ga[i] = (R)ga36[i];
    C s6[6], e6[6];
    for (int c = 0; c < 6; c++) s6[c] = C{(R)s12[2 * c], (R)s12[2 * c + 1]};
    double wr[3], wi[3];
    bool split = false;
    gop::green_point<R>(ga, x, y, kz, nx, ny, nz, s6, e6, wr, wi, split);
    for (int c = 0; c < 6; c++) {
        e12[2 * c] = (double)e6[c].x;
        e12[2 * c + 1] = (double)e6[c].y;
    }
    for (int c = 0; c < 3; c++) {
        w6[c] = wr[c];
        w6[3 + c] = wi[c];
    }
    return split ? 1 : 0;
}
} // namespace

// gop::green_point at half-spectrum entry (x, y, kz) != 0 in FP64 (fp32 = 0)
// or FP32: ga36 the Green coefficients A = C0 : (n x n) as the library builds
// them, s12 the stress transform (11 22 33 23 13 12, real and imaginary
// parts interleaved); e12 the symmetric part of e-hat in the same layout,
// w6 the antisymmetric part (23 13 12 real, then imaginary).  Returns the
// split flag.
extern "C" int gpc_green_point(int fp32, const double *ga36, int x, int y, int kz, int nx, int ny, int nz,
                               const double *s12, double *e12, double *w6)
{
    if (fp32) return green_point_as<float, cplx32>(ga36, x, y, kz, nx, ny, nz, s12, e12, w6);
    return green_point_as<double, cplx64>(ga36, x, y, kz, nx, ny, nz, s12, e12, w6);
}
