// newton.hip -- per-voxel viscoplastic stress update (src/evolution.c:2-310)
// with the power-law rate of src/constitutive.c:5-67, in FP64, one voxel per
// lane, plus the block partial sums of the convergence norms and the stress
// averages that get_smacro needs (src/evolution.c:1415-1437, 473-497).
//
// The arithmetic follows the reference's: the same residual and rate law in
// the same basis, the NR_MODIFIED backtracking line search with its
// constants, branches and the TOLX rejection of small steps.  Rewritten, with
// the same mathematics:
//   * the Newton step: J step = -R with J = I + C0 (S + dt D) is solved as
//     K step = -S0 R with the symmetric positive definite K = S0 + S + dt D
//     (S0 = C0^-1) by Cholesky, and the merit gradient J^T R as K (C0 R) --
//     the reference forms LU_inv_66(J) with pivoting and multiplies;
//   * S = LU_inv_66(C_gr) is a per-orientation constant (read from a table,
//     upper triangle), and d(edot)/d(sigma) is symmetric, so only upper
//     triangles are formed; NR_quality never uses d(edot)/d(sigma);
//   * |x|^n by repeated squaring for integral n, 1/CRSS multiplied instead of
//     divided, Schmid tensors as sym((Rn) x (Rb)), fused multiply-adds.
// A non-positive-definite K returns status 35, the reference's singular-LU
// exit.
//
// Data placement: the 12x5 Schmid tensors of a voxel are derived at kernel
// start from its rotation (72 B read instead of 480 B) and kept in a
// wave-private LDS slab laid out [component][lane] (conflict-free), which
// keeps them out of the register file; C0 and S0 (packed triangles) are a
// 42-entry LDS table read by broadcast; all other per-voxel arrays live in
// registers with compile-time indices only.  One wave per SIMD: the solve
// holds about 420 registers, so the code is shaped for latency (next Schmid
// vector read ahead, constants hoisted, no scalar loops on the hot path).
#include <hip/hip_ext.h>

#include "evpgpu_impl.h"
#include "linesearch.h"

namespace {

__device__ __forceinline__ double ipow(double x, int m)
{
    double r = 1.0, b = x;
    while (m) {
        if (m & 1) r *= b;
        b *= b;
        m >>= 1;
    }
    return r;
}

// x^m for a wave-uniform 0 <= m < 32: the products of ipow in the same
// order, straight-line (a scalar-masked select per exponent bit instead of a
// scalar loop: with one wave per SIMD every scalar instruction and branch
// costs an issue slot)
__device__ __forceinline__ double ipow5(double x, int m)
{
    double r = 1.0, b = x;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const double t = r * b;
        r = ((m >> i) & 1) ? t : r;
        if (i < 4) b *= b;
    }
    return r;
}

// x^M for a compile-time M: the products of ipow with the bit tests folded
// (x^9 is three squarings and one product)
template <int M> __device__ __forceinline__ double ipowc(double x)
{
    double r = 1.0, b = x;
#pragma unroll
    for (int m = M; m; m >>= 1) {
        if (m & 1) r *= b;
        if (m > 1) b *= b;
    }
    return r;
}

// Rate-law evaluation: RM_POW pow() for a non-integral exponent, RM_LOOP
// repeated squaring for an integral one, RM_FAST the common case of every
// crystal phase having EVPGPU_NSYSMX systems with one integral n <= 32 and
// one reference rate (constants hoisted out of the system loop, straight-
// line squaring; for n = 10 and n = 20 the exponent is a template constant).
// RM_DISL: the dislocation-density law StrainRate_Orowan
// (src/constitutive.c:507-721), with its per-voxel constants formed once
// per increment.
enum { RM_POW = 0, RM_LOOP = 1, RM_FAST = 2, RM_DISL = 3 };

// chg_basis opt 2 for a symmetric tensor given as (11 22 33 23 13 12)
__device__ __forceinline__ void sym_to_v6(const double t[6], double v[6])
{
    const double r2 = RSQ2, r3 = RSQ3, r6 = RSQ6;
    v[0] = t[0] * (-1. * r2) + t[1] * r2;
    v[1] = t[0] * (-1. * r6) + t[1] * (-1. * r6) + t[2] * (2. * r6);
    v[2] = t[3] * r2 + t[3] * r2;
    v[3] = t[4] * r2 + t[4] * r2;
    v[4] = t[5] * r2 + t[5] * r2;
    v[5] = t[0] * r3 + t[1] * r3 + t[2] * r3;
}

// chg_basis opt 1 -> symmetric (11 22 33 23 13 12)
__device__ __forceinline__ void v6_to_sym(const double v[6], double t[6])
{
    const double r2 = RSQ2, r3 = RSQ3, r6 = RSQ6;
    t[0] = v[0] * (-1. * r2) + v[1] * (-1. * r6) + v[5] * r3;
    t[1] = v[0] * r2 + v[1] * (-1. * r6) + v[5] * r3;
    t[2] = v[1] * (2. * r6) + v[5] * r3;
    t[3] = v[2] * r2;
    t[4] = v[3] * r2;
    t[5] = v[4] * r2;
}

// 1/d for a positive normal d: the hardware estimate refined by two Newton
// steps y += y (1 - d y) (the relative error e becomes e^2 per step)
__device__ __forceinline__ double rcp_refined(double d)
{
    double y = __builtin_amdgcn_rcp(d);
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const double e = fma(-d, y, 1.0);
        y = fma(y, e, y);
    }
    return y;
}

constexpr int NB_THREADS = 64; // one wave per block: the LDS slab is wave-private

struct RateIn {
    const double *sc;                              // LDS slab, sc[(s*5+j)*64]
    const double *k0;                              // LDS: C0 then S0, packed triangles
    double itf[EVPGPU_NSYSMX], itb[EVPGPU_NSYSMX]; // 1/CRSS forward, backward;
                                                   // RM_DISL: itf the passing stress
    const double *tpn;                             // RM_DISL: this voxel's constants, system
    long long tps;                                 // stride tps: passing stress, 1/(tau_0 +
                                                   // tau_cut), gam_o, first passing stress
};
#define SC(i, j) in.sc[((i)*5 + (j)) * NB_THREADS]

// StrainRate_eval (src/constitutive.c:5-67) as one pass over the slip
// systems: each system's Schmid vector is read from LDS once and feeds the
// resolved shear, the rate, edot and d(edot)/d(sigma).  The sums run over the
// systems in ascending order, as the reference's separate loops do.
// DD: dt d(edot)/d(sigma) is added into the packed Newton matrix kk (which
// holds S0 + S on entry).  GD: the slip rates are kept (gamdot).
template <bool DD, int RM, bool GD, int NE>
__device__ __forceinline__ void srate(const NewtonParams &P, int p, const RateIn &in,
                                      const double s[6], double ed[5], double kk[21],
                                      double gd[EVPGPU_NSYSMX])
{
    // (the per-system test stays a branch on the phase's count even where it
    // is always 12: without it the compiler interleaves the systems'
    // arithmetic and the fast variant spills)
    const int ns = P.nsys[p];
    const int intn = P.int_n[p];
    // RM_FAST: the phase's rate constants, as the general path multiplies
    // them ((gam0 n) p1) itau
    const double g0 = P.gam0[p][0], g0n = P.gam0[p][0] * P.nrs[p][0];
    // RM_DISL: 1/(tau_0 + tau_cut) and gam_o of every system, loaded
    // together before the loop (read at their use, under the per-lane test
    // of the passing stress, each active system waited on its own loads)
    double itcs[EVPGPU_NSYSMX], g0ks[EVPGPU_NSYSMX];
    if (RM == RM_DISL) {
#pragma unroll
        for (int k = 0; k < EVPGPU_NSYSMX; k++) {
            itcs[k] = in.tpn[(EVPGPU_NSYSMX + k) * in.tps];
            g0ks[k] = in.tpn[(2 * EVPGPU_NSYSMX + k) * in.tps];
        }
    }
#pragma unroll
    for (int i = 0; i < 5; i++) ed[i] = 0.;
    double scn[5];
#pragma unroll
    for (int j = 0; j < 5; j++) scn[j] = SC(0, j);
#pragma unroll
    for (int k = 0; k < EVPGPU_NSYSMX; k++) {
        // the next system's Schmid vector is read before this one's
        // arithmetic, so the LDS latency is hidden behind it
        double sc[5];
#pragma unroll
        for (int j = 0; j < 5; j++) sc[j] = scn[j];
        if (k + 1 < EVPGPU_NSYSMX)
#pragma unroll
            for (int j = 0; j < 5; j++) scn[j] = SC(k + 1, j);
        if (k < ns) {
            double r = sc[0] * s[0] + sc[1] * s[1] + sc[2] * s[2] + sc[3] * s[3] + sc[4] * s[4];
            double r1 = 0.0, r2 = 0.0; // the rate and its derivative in the resolved shear
            if (RM == RM_DISL) {
                // gam = gam_o sinh((|rss| - tau_pass)/(tau_0 + tau_cut)) sign(rss)
                // above the passing stress, 0 below (FORWARD_PLUS_BACKWARD)
                const double a = fabs(r);
                if (a > in.itf[k]) {
                    const double itc = itcs[k];
                    const double g0k = g0ks[k];
                    const double x = (a - in.itf[k]) * itc;
                    // sinh and cosh from one expm1: with m = e^x - 1 and
                    // q = e^-x, sinh x = m (1 + q) / 2 (no cancellation at
                    // small x) and cosh x = (1 + m + q) / 2
                    const double m = expm1(x), q = rcp_refined(1.0 + m);
                    r2 = g0k * (0.5 * m * (1.0 + q)) * (double)(2 * (r > 0) - 1);
                    if (DD) r1 = g0k * (0.5 * ((1.0 + m) + q)) * itc;
                }
            } else {
            double itau = (r < 0) ? in.itb[k] : in.itf[k];
            r = r * itau;
            double a = fabs(r), p1, p2;
            if (RM == RM_FAST) {
                p1 = NE > 0 ? ipowc<(NE > 0 ? NE - 1 : 0)>(a) : ipow5(a, intn - 1);
                p2 = p1 * a;
            } else if (RM == RM_LOOP) {
                p1 = ipow(a, intn - 1);
                p2 = p1 * a;
            } else {
                p1 = pow(a, P.nrs[p][k] - 1.);
                p2 = pow(a, P.nrs[p][k]);
            }
            r2 = (RM == RM_FAST ? g0 : P.gam0[p][k]) * p2 * (double)(2 * (r > 0) - 1);
            if (DD) r1 = (RM == RM_FAST ? g0n : P.gam0[p][k] * P.nrs[p][k]) * p1 * itau;
            }
            if (GD) gd[k] = r2;
#pragma unroll
            for (int i = 0; i < 5; i++) ed[i] += sc[i] * r2;
            if (DD) {
                r1 *= P.dt;
#pragma unroll
                for (int i = 0; i < 5; i++) {
                    const double w = sc[i] * r1;
#pragma unroll
                    for (int j = i; j < 5; j++) kk[PK(i, j)] += w * sc[j];
                }
            }
        }
    }
}

// 1/sqrt(d) for a positive normal d: the hardware estimate refined by two
// Newton steps y += y (1 - d y^2) / 2 (the estimate's relative error e
// becomes about 1.5 e^2 per step), to within about an ulp
__device__ __forceinline__ double rsqrt_refined(double d)
{
    double y = __builtin_amdgcn_rsq(d);
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const double e = fma(-(d * y), y, 1.0);
        y = fma(0.5 * y, e, y);
    }
    return y;
}

// Cholesky factorisation of the packed symmetric 6x6 k (in place: the lower
// factor L, stored transposed in the same packing) and the solve
// L L^T x = b (b overwritten).  Returns false if k is not positive definite.
__device__ __forceinline__ bool chol6_solve(double k[21], double b[6])
{
    double inv[6];
#pragma unroll
    for (int j = 0; j < 6; j++) {
        double d = k[PK(j, j)];
#pragma unroll
        for (int q = 0; q < j; q++) d -= k[PK(q, j)] * k[PK(q, j)];
        if (!(d > 0.0)) return false;
        inv[j] = rsqrt_refined(d);
#pragma unroll
        for (int i = j + 1; i < 6; i++) {
            double s = k[PK(j, i)];
#pragma unroll
            for (int q = 0; q < j; q++) s -= k[PK(q, i)] * k[PK(q, j)];
            k[PK(j, i)] = s * inv[j];
        }
    }
#pragma unroll
    for (int i = 0; i < 6; i++) {
        double s = b[i];
#pragma unroll
        for (int q = 0; q < i; q++) s -= k[PK(q, i)] * b[q];
        b[i] = s * inv[i];
    }
#pragma unroll
    for (int i = 5; i >= 0; i--) {
        double s = b[i];
#pragma unroll
        for (int q = i + 1; q < 6; q++) s -= k[PK(i, q)] * b[q];
        b[i] = s * inv[i];
    }
    return true;
}

// An empty asm statement that takes x in a register: the value must be
// ready at this point (its load is not moved below it)
__device__ __forceinline__ void keep_reg(const double &x)
{
    __asm__ volatile("" ::"v"(x));
}

// Per-voxel constants of the Newton solve: the compliance S = LU_inv_66(C_gr)
// (symmetric; its upper triangle), lambda and eps_p - E in the B-basis.
struct Vox {
    double S[21];
    double xl[6], bp[6];
    double signorm; // Frobenius norm of lambda (src/evolution.c:92-96)
};

// R(s) = s - lambda + C0 t,  t = eps_p + dt edot(s) + S s - E
// (src/evolution.c:126-140; t is the strain mismatch of the error norm)
__device__ __forceinline__ void residual(const NewtonParams &P, const RateIn &in, const Vox &V,
                                         const double sig[6],
                                         const double ed[5], double t[6], double res[6])
{
#pragma unroll
    for (int i = 0; i < 6; i++) {
        double a = V.bp[i] + (i < 5 ? ed[i] * P.dt : 0.0);
#pragma unroll
        for (int j = 0; j < 6; j++) a += V.S[PK(i, j)] * sig[j];
        t[i] = a;
    }
#pragma unroll
    for (int i = 0; i < 6; i++) {
        double r = sig[i] - V.xl[i];
#pragma unroll
        for (int j = 0; j < 6; j++) r += in.k0[PK(i, j)] * t[j];
        res[i] = r;
    }
}

template <int RM, bool GD, int NE>
__device__ __forceinline__ double nr_quality(const NewtonParams &P, int p, const RateIn &in,
                                             const Vox &V, const double sig[6],
                                             double gd[EVPGPU_NSYSMX])
{
    double ed[5], t[6], res[6];
    srate<false, RM, GD, NE>(P, p, in, sig, ed, nullptr, gd);
    residual(P, in, V, sig, ed, t, res);
    double f = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++) f += res[i] * res[i];
    return f / 2.0;
}

// returns 0, EVPGPU_ERR_SINGULAR or EVPGPU_ERR_LINESEARCH
// FAST: an iteration whose step the reference's line search is certain to
// reject (TOLX) ends right after the residual.  Since D >= 0, K = S0 + S +
// dt D >= S0 + S, so |step|_inf <= |S0 R|_2 / lambda_min(S0 + S); when that
// bound is below TOLX (P.kmin_tolx2 holds (0.99 TOLX lambda_min)^2, the
// minimum over all orientations) the reference's test
// max|step_i|/max(|s_i|,1) < TOLX holds and its first trial is rejected,
// leaving s, edot, the error norms and the
// counts exactly as computed here.  Only gamdot (from the rejected trial
// point) differs; the loop recomputes the last Newton launch of an increment
// with FAST off before gamdot is used, so FAST launches do not form gamdot.
template <int RM, bool FAST, int NE>
__device__ int update_voxel(const NewtonParams &P, int p, RateIn &in, const Vox &V,
                            double sig6[6], double edotp[6], double gd[EVPGPU_NSYSMX],
                            double &conv_s, double &conv_e, int &iterl, int &trials)
{
    constexpr bool GD = !FAST;
    trials = 0;
    double ed[5], res[6], old[6], k[21];
    double ce = 0.0;
    iterl = 0;
    double erral = 10 * P.nr_tol;
#pragma unroll
    for (int i = 0; i < 6; i++) sig6[i] = V.xl[i];
    while ((iterl < 500) && (fabs(erral) > fabs(P.nr_tol))) {
        // read C0, S0 and the Schmid vectors from LDS where they are used:
        // hoisted out of the loop they would occupy 204 more registers
        __asm__ volatile("" ::: "memory");
        iterl++;
#pragma unroll
        for (int i = 0; i < 6; i++) old[i] = sig6[i];
        // the Newton matrix K = S0 + S + dt D (below), D added by srate
#pragma unroll
        for (int i = 0; i < 6; i++)
#pragma unroll
            for (int j = i; j < 6; j++) k[PK(i, j)] = in.k0[21 + PK(i, j)] + V.S[PK(i, j)];
        srate<true, RM, GD, NE>(P, p, in, sig6, ed, k, gd);
        if (RM == RM_DISL && P.dd_first && iterl == 1) {
            // the first call used the previous call's mobile density; every
            // later one sees the density that call formed
#pragma unroll
            for (int s = 0; s < EVPGPU_NSYSMX; s++)
                if (s < P.nsys[p]) in.itf[s] = in.tpn[s * in.tps];
        }
#pragma unroll
        for (int i = 0; i < 6; i++) edotp[i] = i < 5 ? ed[i] : 0.0;
        double gradf[6], step[6];
        {
            double t[6];
            residual(P, in, V, sig6, ed, t, res);
            ce = 0.0;
#pragma unroll
            for (int i = 0; i < 6; i++) ce += t[i] * t[i];
        }
        // S0 R: the right-hand side of the step, and its norm bounds the step
#pragma unroll
        for (int i = 0; i < 6; i++) {
            double b = 0.0;
#pragma unroll
            for (int j = 0; j < 6; j++) b += in.k0[21 + PK(i, j)] * res[j];
            step[i] = b;
        }
        bool skip = false;
        if (FAST) {
            double y2 = 0.0;
#pragma unroll
            for (int i = 0; i < 6; i++) y2 += step[i] * step[i];
            skip = (y2 > 0.0) && (y2 < P.kmin_tolx2);
        }
        if (skip) {
            trials++; // the reference's one rejected trial
        } else {
        {
            // The Newton matrix J = I + C0 (S + dt D) times S0 = C0^-1 is the
            // symmetric positive definite K = S0 + S + dt D: the step solves
            // K step = -S0 R (Cholesky), and the merit-function gradient
            // J^T R = R + (S + dt D) C0 R equals K (C0 R).
            double u[6];
#pragma unroll
            for (int i = 0; i < 6; i++) {
                double a = 0.0;
#pragma unroll
                for (int j = 0; j < 6; j++) a += in.k0[PK(i, j)] * res[j];
                u[i] = a;
            }
#pragma unroll
            for (int i = 0; i < 6; i++) {
                double g = 0.0;
#pragma unroll
                for (int j = 0; j < 6; j++) g += k[PK(i, j)] * u[j];
                gradf[i] = g;
            }
            if (!chol6_solve(k, step)) return EVPGPU_ERR_SINGULAR;
#pragma unroll
            for (int i = 0; i < 6; i++) step[i] = -step[i];
        }
        double f_old = 0.0;
#pragma unroll
        for (int i = 0; i < 6; i++) f_old += res[i] * res[i];
        f_old *= 0.5;
        const double stpmax = 10.0, ALF = 1.E-4, TOLX = P.tolx;
        double sum = 0.0, slope = 0.0, test = 0.0;
#pragma unroll
        for (int i = 0; i < 6; i++) sum += step[i] * step[i];
        sum = sqrt(sum);
        if (sum > stpmax) {
#pragma unroll
            for (int i = 0; i < 6; i++) step[i] *= stpmax / sum;
        }
#pragma unroll
        for (int i = 0; i < 6; i++) slope += gradf[i] * step[i];
        if (slope >= 0) return EVPGPU_ERR_LINESEARCH;
#pragma unroll
        for (int i = 0; i < 6; i++) {
            double d = fabs(old[i]);
            double t = fabs(step[i]) / (d > 1.0 ? d : 1.0);
            if (t > test) test = t;
        }
        double alamin = TOLX / test, alam = 1.0, alam2 = 0.0, f2 = 0.0;
        // The first trial (alam = 1) is rejected when 1 < alamin, whatever
        // its merit value: the step is left undone and only gamdot (from the
        // trial point) is affected, which FAST launches do not form -- so
        // they skip the evaluation.  This is how nearly every voxel's solve
        // ends: an accepted step has relative size >= TOLX, which keeps the
        // relative-step test (1e-10) open, so the last step is a TOLX reject.
        if (FAST && alam < alamin) {
            trials++;
        } else {
        while (true) {
            __asm__ volatile("" ::: "memory");
#pragma unroll
            for (int i = 0; i < 6; i++) sig6[i] = old[i] + alam * step[i];
            double f_new = nr_quality<RM, GD, NE>(P, p, in, V, sig6, gd);
            trials++;
            if (alam < alamin) {
#pragma unroll
                for (int i = 0; i < 6; i++) sig6[i] = old[i];
                break;
            } else if (f_new <= f_old + ALF * alam * slope) {
                break;
            }
            const double next = ls_backtrack(alam, alam2, f_new, f2, f_old, slope);
            alam2 = alam;
            f2 = f_new; // the backtrack value the cubic step needs (Q18)
            alam = next;
        }
        } // line search
        } // !skip
        double conv_nr = 0.0;
#pragma unroll
        for (int i = 0; i < 6; i++) conv_nr += (sig6[i] - old[i]) * (sig6[i] - old[i]);
        erral = sqrt(conv_nr) / V.signorm;
    }
    // the norms of the last iteration (the reference overwrites them in
    // every iteration)
    double cs = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++) cs += (sig6[i] - V.xl[i]) * (sig6[i] - V.xl[i]);
    conv_s = sqrt(cs);
    conv_e = sqrt(ce);
    return 0;
}

// One lane per voxel.  Reads lambda from sig_in and writes the new stress to
// sig (separate buffers), edot and -- with FAST off -- gamdot; block partials
// (NRED doubles per block) of err_s, err_e, Newton iterations, sum sig, sum
// sig over phase 1 and line-search trials.
template <int RM, bool CSYM, bool FAST, int NE>
__global__ __launch_bounds__(NB_THREADS) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_newton(
    NewtonParams P, long long N, const int *__restrict__ phase, const int *__restrict__ orient,
    const double *__restrict__ sgr, const double *sig_in, double *sig, const double *__restrict__ E,
    const double *__restrict__ epsp, double *__restrict__ edot, const double *__restrict__ tran,
    const double *__restrict__ crss, double *__restrict__ gamdot, double *__restrict__ partials,
    int *__restrict__ status, double *__restrict__ err_vox, int *__restrict__ niter_vox)
{
    if (P.skip && *P.skip) return;
    __shared__ double slab[EVPGPU_NSYSMX * 5 * NB_THREADS];
    // C0 and S0 (packed upper triangles) as broadcast LDS reads: kept in
    // scalar registers they do not fit, and each use of a spilled one costs
    // two lane reads
    __shared__ double k0[42];
    long long v = (long long)blockIdx.x * NB_THREADS + threadIdx.x;
    // the voxel's phase and orientation index go out with the C0/S0 loads,
    // unconditionally (a lane past the end reads the last voxel's and drops
    // them), so that one wait covers all four: with one wave per SIMD every
    // wait on memory stalls the SIMD
    const long long vc = v < N ? v : N - 1;
    const int ph_raw = phase[vc], ori = orient[vc];
    const double c0v = P.c0p[threadIdx.x % 21], s0v = P.s0p[threadIdx.x % 21];
    const int ph = v < N ? ph_raw : 1;
    if (threadIdx.x < 21) {
        k0[threadIdx.x] = c0v;
        k0[21 + threadIdx.x] = s0v;
    }
    __syncthreads();
    // this lane's err_s, err_e, Newton iterations and line-search trials
    double r_cs = 0.0, r_ce = 0.0, r_it = 0.0, r_tr = 0.0;
    int pl = 0;
    if (v < N) {
        pl = ph - 1;
        pl = pl < 0 ? 0 : (pl > 1 ? 1 : pl);
    }
    // the two phases' gas flags as a select of kernel arguments, not an
    // indexed (memory) read
    const bool gas_l = pl == 0 ? P.is_gas[0] != 0 : P.is_gas[1] != 0;
    // Crystal voxels, one phase of the wave at a time: inside the loop the
    // phase index p is wave-uniform, so the rate-law constants, slip-system
    // counts and crystal axes are scalar operands and the slip-system loops
    // branch on scalars (a wave holds one crystal phase almost always).
    bool pending = v < N && !gas_l;
    // every pass clears at least the lead lane, so it ends after one pass
    // per distinct phase (two at most)
    for (unsigned long long todo = __ballot(pending); todo; todo = __ballot(pending)) {
        const int p = __builtin_amdgcn_readlane(pl, __ffsll(todo) - 1);
        // lanes of phase p, tested through the ballot mask: a branch on
        // pl == p itself would let the compiler substitute the per-lane pl
        // for the uniform p inside the branch
        const unsigned long long sel = __ballot(pending && pl == p);
        if ((sel >> (threadIdx.x & 63)) & 1ull) {
            pending = false;
            Vox V;
            RateIn in;
            // every load of the voxel's inputs first -- stress, plastic and
            // total strain, the orientation's S, the rotation and the
            // per-system CRSS (all 12: the arrays hold them whatever the
            // phase's count) -- then the arithmetic: one wait on memory
            // instead of one per group
            double t[6], tp[6], te[6], T9[9], c1[EVPGPU_NSYSMX], c2[EVPGPU_NSYSMX];
#pragma unroll
            for (int c = 0; c < 6; c++) t[c] = sig_in[c * N + v];
#pragma unroll
            for (int c = 0; c < 6; c++) tp[c] = epsp[c * N + v];
#pragma unroll
            for (int c = 0; c < 6; c++) te[c] = E[c * N + v];
            const double *sv = sgr + 36LL * ori;
#pragma unroll
            for (int i = 0; i < 6; i++)
#pragma unroll
                for (int j = i; j < 6; j++) V.S[PK(i, j)] = sv[i * 6 + j];
#pragma unroll
            for (int q = 0; q < 9; q++) T9[q] = tran[q * N + v];
            const long long S12 = (long long)EVPGPU_NSYSMX * N;
#pragma unroll
            for (int s = 0; s < EVPGPU_NSYSMX; s++) {
                if (RM == RM_DISL) {
                    c1[s] = crss[(P.dd_first ? 3 * S12 : 0) + s * N + v];
                } else {
                    c1[s] = crss[(s * 2) * N + v];
                    c2[s] = CSYM ? c1[s] : crss[(s * 2 + 1) * N + v];
                }
            }
            // |lambda|_F over the full 3x3, row-major order
            V.signorm = sqrt(t[0] * t[0] + t[5] * t[5] + t[4] * t[4] + t[5] * t[5] + t[1] * t[1] +
                             t[3] * t[3] + t[4] * t[4] + t[3] * t[3] + t[2] * t[2]);
            sym_to_v6(t, V.xl);
            {
                double e6[6], s6[6];
                sym_to_v6(tp, e6);
                sym_to_v6(te, s6);
#pragma unroll
                for (int i = 0; i < 6; i++) V.bp[i] = e6[i] - s6[i];
            }
            // the fast variants have every crystal phase on 12 systems: a
            // compile-time count keeps the per-system loops free of branches
            const int ns = RM == RM_FAST || RM == RM_DISL ? EVPGPU_NSYSMX : P.nsys[p];
            // Schmid tensors in sample axes: sym((T n) x (T b)) in the 5-D basis
            // (update_schmid, src/kinematics.c:430-463), into this lane's LDS column
            double T[3][3];
#pragma unroll
            for (int q = 0; q < 9; q++) T[q / 3][q % 3] = T9[q];
            in.sc = slab + threadIdx.x;
            in.k0 = k0;
            // unrolled (straight-line where the count is the constant 12), so
            // that the loads above stay ahead of this arithmetic
#pragma unroll
            for (int s = 0; s < EVPGPU_NSYSMX; s++) {
                if (s >= ns) break;
                double tn[3], tb[3];
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    tn[a] = T[a][0] * P.bn[p][s][0] + T[a][1] * P.bn[p][s][1] + T[a][2] * P.bn[p][s][2];
                    tb[a] = T[a][0] * P.bb[p][s][0] + T[a][1] * P.bb[p][s][1] + T[a][2] * P.bb[p][s][2];
                }
                double P00 = tn[0] * tb[0], P11 = tn[1] * tb[1], P22 = tn[2] * tb[2];
                double P12 = (tn[1] * tb[2] + tn[2] * tb[1]) / 2.0;
                double P02 = (tn[0] * tb[2] + tn[2] * tb[0]) / 2.0;
                double P01 = (tn[0] * tb[1] + tn[1] * tb[0]) / 2.0;
                double *col = slab + threadIdx.x + s * 5 * NB_THREADS;
                col[0 * NB_THREADS] = P00 * (-1. * RSQ2) + P11 * RSQ2;
                col[1 * NB_THREADS] = P00 * (-1. * RSQ6) + P11 * (-1. * RSQ6) + P22 * (2. * RSQ6);
                col[2 * NB_THREADS] = P12 * RSQ2 + P12 * RSQ2;
                col[3 * NB_THREADS] = P02 * RSQ2 + P02 * RSQ2;
                col[4 * NB_THREADS] = P01 * RSQ2 + P01 * RSQ2;
            }
            // the voxel's other inputs are taken here, after the Schmid
            // arithmetic: left alone, the compiler sinks their loads past the
            // Newton loop's entry test and the wave waits on memory twice
#pragma unroll
            for (int i = 0; i < 6; i++) keep_reg(V.bp[i]);
#pragma unroll
            for (int i = 0; i < 21; i++) keep_reg(V.S[i]);
#pragma unroll
            for (int k = 0; k < EVPGPU_NSYSMX; k++) {
                keep_reg(c1[k]);
                if (!CSYM && RM != RM_DISL) keep_reg(c2[k]);
            }
            if (RM == RM_DISL) {
                // crss holds the increment's constants: passing stress,
                // 1/(tau_0 + tau_cut), gam_o, the first call's passing stress
                in.tpn = crss + v;
                in.tps = N;
#pragma unroll
                for (int s = 0; s < EVPGPU_NSYSMX; s++) in.itf[s] = s < ns ? c1[s] : 0.0;
            } else {
#pragma unroll
            for (int s = 0; s < EVPGPU_NSYSMX; s++) {
                if (s < ns) {
                    in.itf[s] = 1.0 / c1[s];
                    in.itb[s] = CSYM ? in.itf[s] : 1.0 / c2[s];
                } else {
                    in.itf[s] = in.itb[s] = 1.0;
                }
            }
            }
            double s6[6], e6[6], gd[EVPGPU_NSYSMX], cs, ce;
            int it, ntr;
            int st = update_voxel<RM, FAST, NE>(P, p, in, V, s6, e6, gd, cs, ce, it, ntr);
            if (st) {
                atomicCAS(status, 0, st);
            } else {
                v6_to_sym(s6, t);
#pragma unroll
                for (int c = 0; c < 6; c++) sig[c * N + v] = t[c];
                v6_to_sym(e6, t);
#pragma unroll
                for (int c = 0; c < 6; c++) edot[c * N + v] = t[c];
                if (!FAST)
#pragma unroll
                    for (int s = 0; s < EVPGPU_NSYSMX; s++)
                        if (s < ns) gamdot[s * N + v] = gd[s];
                r_cs = cs;
                r_ce = ce;
                r_it = (double)it;
                r_tr = (double)ntr;
                if (err_vox) {
                    err_vox[2 * v] = cs;
                    err_vox[2 * v + 1] = ce;
                }
                if (niter_vox) niter_vox[v] = it;
            }
        }
    }
    double red[NRED];
#pragma unroll
    for (int i = 0; i < NRED; i++) red[i] = 0.0;
    red[0] = r_cs;
    red[1] = r_ce;
    red[2] = r_it;
    red[15] = r_tr;
    if (v < N) {
        if (gas_l) {
#pragma unroll
            for (int c = 0; c < 6; c++) {
                sig[c * N + v] = 0.0;
                edot[c * N + v] = 0.0;
            }
            if (err_vox) err_vox[2 * v] = err_vox[2 * v + 1] = 0.0;
            if (niter_vox) niter_vox[v] = 0;
        }
#pragma unroll
        for (int c = 0; c < 6; c++) {
            double sc = sig[c * N + v];
            red[3 + c] = sc;
            red[9 + c] = (phase[v] == 1) ? sc : 0.0;
        }
    }
    // Block partial sums, transposed through the slab (no longer needed):
    // value i is summed by lanes 4i..4i+3 over 16 entries each, then across
    // the four (a + b == b + a, so all four hold the same sum): a fixed order
    // at a quarter of the instructions of 16 butterfly reductions.
    static_assert(NB_THREADS == 64 && NRED * 4 == NB_THREADS && NRED * 65 <= EVPGPU_NSYSMX * 5 * NB_THREADS,
                  "one wave per block, 16 partial sums");
    const int lane = threadIdx.x;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NRED; i++) slab[i * 65 + lane] = red[i];
    __syncthreads();
    const int vi = lane >> 2, q = lane & 3;
    double sum = 0.0;
#pragma unroll
    for (int j = 0; j < 16; j++) sum += slab[vi * 65 + q * 16 + j];
    sum += __shfl_xor(sum, 1, 64);
    sum += __shfl_xor(sum, 2, 64);
    if (q == 0) partials[(long long)blockIdx.x * NRED + vi] = sum;
}

// Deterministic two-stage sum of nblk rows of nval (<= 16) partials.
// Stage 1: workgroup g sums rows [g*per, (g+1)*per): thread t owns value
// t % 16 and rows t/16, t/16 + 16, ... (coalesced row-major reads), then the
// 16 row lanes of each value are summed in LDS in a fixed order.
// Stage 2: one workgroup sums the stage-1 rows the same way.
constexpr int RED_MAXV = 16;

// status (stage 2 only, may be null): the Newton status word, appended as
// out[nval] so that one copy brings back the sums and the status.  skip (may
// be null): the launch does nothing when *skip is set.
__global__ __launch_bounds__(256) void k_reduce_rows(const double *__restrict__ partials, int nblk,
                                                     int nval, int per, double *__restrict__ out,
                                                     const int *__restrict__ status,
                                                     const int *__restrict__ skip)
{
    if (skip && *skip) return;
    if (status && blockIdx.x == 0 && threadIdx.x == 0) out[nval] = (double)*status;
    __shared__ double sh[256];
    const int v = threadIdx.x % RED_MAXV, r0 = threadIdx.x / RED_MAXV;
    const long long b0 = (long long)blockIdx.x * per;
    const long long b1 = min((long long)nblk, b0 + per);
    double s = 0.0;
    if (v < nval) {
        // unrolled so that the loads go out together (the sum keeps its order)
#pragma unroll 4
        for (long long b = b0 + r0; b < b1; b += 256 / RED_MAXV) s += partials[b * nval + v];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < RED_MAXV) {
        double t = 0.0;
        for (int r = 0; r < 256 / RED_MAXV; r++) t += sh[r * RED_MAXV + threadIdx.x];
        if (threadIdx.x < nval) out[(long long)blockIdx.x * nval + threadIdx.x] = t;
    }
}

} // namespace

int newton_launch(evpgpu_ctx *c, const NewtonParams &P, const double *sig_in, double *sig, const double *E,
                  const double *epsp, double *edot, const double *tran, const double *crss,
                  double *gamdot, const int *phase, const int *orient, const double *sgr,
                  long long n, double *partials, int *nblk_out, int *status, double *err_vox,
                  int *niter_vox, int fast, hipEvent_t ev_start, hipEvent_t ev_stop)
{
    int nblk = (int)((n + NB_THREADS - 1) / NB_THREADS);
    // the integral-exponent variants need every crystal phase to have one;
    // the fast one also n <= 32, a full set of systems and one gam0
    bool intn = true, fastrm = true;
    for (int q = 0; q < 2; q++) {
        if (P.is_gas[q]) continue;
        if (P.int_n[q] == 0) intn = false;
        bool one = P.int_n[q] >= 1 && P.int_n[q] <= 32 && P.nsys[q] == EVPGPU_NSYSMX;
        for (int k = 0; k < P.nsys[q]; k++)
            if (P.gam0[q][k] != P.gam0[q][0] || P.nrs[q][k] != P.nrs[q][0]) one = false;
        if (!one) fastrm = false;
    }
    const int rm = !intn ? RM_POW : (fastrm ? RM_FAST : RM_LOOP);
    // the dislocation-density variant assumes 12 systems in every crystal phase
    if (P.dd)
        for (int q = 0; q < 2; q++)
            if (!P.is_gas[q] && P.nsys[q] != EVPGPU_NSYSMX) return EVPGPU_ERR_ARG;
    // compile-time exponents for the common n = 10 and n = 20 (every crystal
    // phase the same n, forward = backward CRSS)
    int ne = 0;
    if (rm == RM_FAST && P.crss_sym) {
        int n0 = 0;
        bool same = true;
        for (int q = 0; q < 2; q++) {
            if (P.is_gas[q]) continue;
            if (n0 && P.int_n[q] != n0) same = false;
            n0 = P.int_n[q];
        }
        if (same && (n0 == 10 || n0 == 20)) ne = n0;
    }
    // the events (when given) take the kernel's own start and end times: no
    // marker packets between the launches
#define NEWTON_ARGS dim3(nblk), dim3(NB_THREADS), 0, c->stream, ev_start, ev_stop, 0, P, n, phase, orient, sgr, \
                    sig_in, sig, E, epsp, edot, tran, crss, gamdot, partials, status, err_vox, niter_vox
#define LAUNCH(A, B, F) hipExtLaunchKernelGGL((k_newton<A, B, F, 0>), NEWTON_ARGS)
#define LAUNCH_NE(E, F) hipExtLaunchKernelGGL((k_newton<RM_FAST, true, F, E>), NEWTON_ARGS)
#define LAUNCH_RM(B, F)                                                                  \
    do {                                                                                 \
        if (ne == 10) LAUNCH_NE(10, F);                                                  \
        else if (ne == 20) LAUNCH_NE(20, F);                                             \
        else if (rm == RM_FAST) LAUNCH(RM_FAST, B, F);                                   \
        else if (rm == RM_LOOP) LAUNCH(RM_LOOP, B, F);                                   \
        else LAUNCH(RM_POW, B, F);                                                       \
    } while (0)
    if (P.dd) {
        if (fast) LAUNCH(RM_DISL, false, true);
        else LAUNCH(RM_DISL, false, false);
    } else if (fast) {
        if (P.crss_sym) LAUNCH_RM(true, true);
        else LAUNCH_RM(false, true);
    } else {
        if (P.crss_sym) LAUNCH_RM(true, false);
        else LAUNCH_RM(false, false);
    }
#undef LAUNCH_RM
#undef LAUNCH_NE
#undef LAUNCH
#undef NEWTON_ARGS
    if (nblk_out) *nblk_out = nblk;
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

// out[nval] = column sums of partials[nblk][nval]; uses c->red2 (REDUCE_G rows)
int reduce_partials(evpgpu_ctx *c, const double *partials, int nblk, int nval, double *out, const int *status)
{
    if (nval > RED_MAXV) return EVPGPU_ERR_ARG;
    int g = nblk < REDUCE_G ? nblk : REDUCE_G;
    int per = (nblk + g - 1) / g;
    g = (nblk + per - 1) / per;
    hipLaunchKernelGGL(k_reduce_rows, dim3(g), dim3(256), 0, c->stream, partials, nblk, nval, per, c->red2,
                       (const int *)nullptr, (const int *)nullptr);
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(256), 0, c->stream, (const double *)c->red2, g, nval, g,
                       out, status, (const int *)nullptr);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

// the first stage of reduce_partials alone: c->red2 = g rows of nval sums
// (the second stage's input, summed the same way by the loop control)
int reduce_stage1(evpgpu_ctx *c, const double *partials, int nblk, int nval, const int *skip, int *g_out)
{
    if (nval > RED_MAXV) return EVPGPU_ERR_ARG;
    int g = nblk < REDUCE_G ? nblk : REDUCE_G;
    int per = (nblk + g - 1) / g;
    g = (nblk + per - 1) / per;
    hipLaunchKernelGGL(k_reduce_rows, dim3(g), dim3(256), 0, c->stream, partials, nblk, nval, per, c->red2,
                       (const int *)nullptr, skip);
    *g_out = g;
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}
