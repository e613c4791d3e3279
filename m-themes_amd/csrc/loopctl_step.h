// loopctl_step.h -- the Green loop's control block and the arithmetic of
// its per-iteration step: get_smacro (src/evolution.c:473-547) on the
// reduced sums and the loop test (src/evolution.c:1354).  Shared by the
// one-workgroup control kernel (loopctl.hip) and a host build that checks
// the step against the oracle's get_smacro bit for bit; both are compiled
// without contraction (-ffp-contract=off), with IEEE division and square root.
#pragma once

#include <math.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define LCHD __host__ __device__ __forceinline__
#else
#define LCHD static inline
#endif

// Lebensohn-Tome basis constants (src/kinematics.c:22-24)
#define RSQ2 0.70710678118654744
#define RSQ3 0.57735026918962584
#define RSQ6 0.40824829046386304

// The Green loop's control, kept on the device so that the host can queue
// iterations ahead of it: after each iteration's Newton reduction a
// one-workgroup kernel finishes get_smacro, the error norms and the stopping
// test (src/evolution.c:1415-1458), and sets `done` when the loop phase ends;
// the kernels of the iterations queued after that see `done` and return.
struct LoopFlags {
    int iter;    // Green iterations carried out in the increment
    int done;    // the phase has ended (converged, at its limit, or failed)
    int status;  // EVPGPU_OK or the error that ended it
    int pad;
};
struct LoopCtl {
    LoopFlags f;
    int limit;   // the phase runs while iter < limit ...
    int conv;    // ... and, when conv, while max(|err_s|, |err_e|) > tol
    double tol, e_vm;
    double dD[9], dD_acum[9], dD_applied[9];
    double add[6];              // the symmetric part of dD, as the strain update adds it
    double sig_avg[9], sig_avg1[9], s_vm, s_vm1;
    double err_s, err_e;
    double nit, ntr;            // Newton iterations and line-search trials, summed
};
// the loading constants get_smacro reads
struct CtlConst {
    double s0[81], scauchy[9];
    double wgt_ph1, scale;
    int velgrad_bc[6], stress_bc[6];
};
static_assert(sizeof(LoopCtl) % 8 == 0 && sizeof(CtlConst) % 8 == 0, "control blocks are whole doubles");

namespace lctl {

constexpr int IJV0[6] = {0, 1, 2, 1, 0, 0};
constexpr int IJV1[6] = {0, 1, 2, 2, 2, 1};

// the basis of chg_basis (src/kinematics.c:14-174): B[i][a][b]
LCHD double basis_entry(int i, int a, int b)
{
    if (i == 0) return a == 0 && b == 0 ? -1. * RSQ2 : (a == 1 && b == 1 ? RSQ2 : 0.0);
    if (i == 1) return a == b ? (a == 2 ? 2. * RSQ6 : -1. * RSQ6) : 0.0;
    if (i == 2) return (a == 1 && b == 2) || (a == 2 && b == 1) ? RSQ2 : 0.0;
    if (i == 3) return (a == 0 && b == 2) || (a == 2 && b == 0) ? RSQ2 : 0.0;
    if (i == 4) return (a == 0 && b == 1) || (a == 1 && b == 0) ? RSQ2 : 0.0;
    return a == b ? RSQ3 : 0.0;
}

// chg_basis opt 2 (first n components)
LCHD void t2_to_v(const double t[9], double *v, int n)
{
#pragma unroll
    for (int i = 0; i < 6; i++) {
        if (i >= n) break;
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) s += t[a * 3 + b] * basis_entry(i, a, b);
        v[i] = s;
    }
}

// chg_basis opt 1 (first n components)
LCHD void v_to_t2(const double *v, double t[9], int n)
{
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < 6; i++) {
                if (i >= n) break;
                s += v[i] * basis_entry(i, a, b);
            }
            t[a * 3 + b] = s;
        }
}

LCHD void sym6_to_full(const double s[6], double t[9])
{
    t[0] = s[0]; t[4] = s[1]; t[8] = s[2];
    t[5] = t[7] = s[3];
    t[2] = t[6] = s[4];
    t[1] = t[3] = s[5];
}

// get_smacro (src/evolution.c:473-547) on the reduced sums
LCHD void smacro(const CtlConst &K, LoopCtl &L, const double *red)
{
    const double wgt = K.scale;
    double s[6], s1[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        s[k] = red[3 + k] * wgt;
        s1[k] = red[9 + k] * (wgt / K.wgt_ph1);
    }
    sym6_to_full(s, L.sig_avg);
    sym6_to_full(s1, L.sig_avg1);
    double dD[9];
#pragma unroll
    for (int i = 0; i < 9; i++) dD[i] = L.dD[i];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const int ii = IJV0[i], jj = IJV1[i];
        double d = 0.0;
        if (K.velgrad_bc[i] == 0) {
#pragma unroll
            for (int k = 0; k < 6; k++) {
                const int kk = IJV0[k], ll = IJV1[k];
                d += K.s0[((ii * 3 + jj) * 3 + kk) * 3 + ll] * K.stress_bc[k] *
                     (K.scauchy[kk * 3 + ll] - L.sig_avg[kk * 3 + ll]);
            }
        }
        dD[ii * 3 + jj] = d;
    }
#pragma unroll
    for (int i = 0; i < 9; i++) {
        L.dD[i] = dD[i];
        L.dD_acum[i] += dD[i];
    }
    double v6[6], dev[9];
    t2_to_v(L.sig_avg, v6, 6);
    v_to_t2(v6, dev, 5);
    double tr = dev[0] + dev[4] + dev[8], a = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double d = dev[i * 3 + j] - (i == j ? tr / 3.0 : 0.0);
            a += (i == j) ? d * d : dev[i * 3 + j] * dev[i * 3 + j];
        }
    L.s_vm = sqrt(3. / 2. * a);
    t2_to_v(L.sig_avg1, v6, 6);
    v_to_t2(v6, dev, 5);
    a = 0.0;
#pragma unroll
    for (int i = 0; i < 9; i++) a += dev[i] * dev[i];
    L.s_vm1 = sqrt(3. / 2. * a);
}

LCHD int phase_done(const LoopCtl &L)
{
    return !(L.f.iter < L.limit && (!L.conv || fmax(fabs(L.err_s), fabs(L.err_e)) > L.tol));
}

} // namespace lctl
