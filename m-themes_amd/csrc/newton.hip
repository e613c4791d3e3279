// newton.hip -- per-voxel viscoplastic stress update (src/evolution.c:2-310)
// with the power-law rate of src/constitutive.c:5-67, in FP64, one voxel per
// lane, plus the block partial sums of the convergence norms and the stress
// averages that get_smacro needs (src/evolution.c:1415-1437, 473-497).
//
// The arithmetic follows the reference's operation order (basis changes with
// their nonzero terms in the same order, Crout LU with implicit pivoting and
// the same 1e-5 thresholds, explicit 6x6 inverse, NR_MODIFIED backtracking
// line search).  Differences that are exact rewrites:
//   * the 5x5 block of d(edot)/d(sigma) is symmetric -- (a*b)*c == (b*a)*c --
//     so only its upper triangle is summed;
//   * NR_quality never uses d(edot)/d(sigma), so it is not formed there;
//   * zero rows/columns of d(edot)/d(sigma) (6th component) are skipped;
//   * LU_inv_66(C_gr) is a per-orientation constant, read from a table.
// Differences in rounding: |x|^n uses repeated squaring for integral n, and
// the compiler fuses multiply-adds.
//
// Registers: all per-voxel arrays are indexed with compile-time indices only
// (row swaps of the pivoting are selects), so nothing goes to scratch.
#include "evpgpu_impl.h"

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

struct RateIn {
    double sc[EVPGPU_NSYSMX][5];
    double tf[EVPGPU_NSYSMX], tb[EVPGPU_NSYSMX];
};

template <bool DD, bool INTN>
__device__ __forceinline__ void srate(const NewtonParams &P, int p, const RateIn &in,
                                      const double s[6], double ed[5], double dd[5][5],
                                      double gd[EVPGPU_NSYSMX])
{
    double r1[EVPGPU_NSYSMX], r2[EVPGPU_NSYSMX];
    const int ns = P.nsys[p];
    const int intn = P.int_n[p];
#pragma unroll
    for (int i = 0; i < EVPGPU_NSYSMX; i++) {
        r1[i] = 0.0;
        r2[i] = 0.0;
        if (i < ns) {
            double r = in.sc[i][0] * s[0] + in.sc[i][1] * s[1] + in.sc[i][2] * s[2] +
                       in.sc[i][3] * s[3] + in.sc[i][4] * s[4];
            double tau = (r < 0) ? in.tb[i] : in.tf[i];
            r = r / tau;
            double a = fabs(r), p1, p2;
            if (INTN) {
                p1 = ipow(a, intn - 1);
                p2 = p1 * a;
            } else {
                p1 = pow(a, P.nrs[p][i] - 1.);
                p2 = pow(a, P.nrs[p][i]);
            }
            r1[i] = P.gam0[p][i] * P.nrs[p][i] * p1 / tau;
            r2[i] = P.gam0[p][i] * p2 * (double)(2 * (r > 0) - 1);
            gd[i] = r2[i];
        }
    }
#pragma unroll
    for (int i = 0; i < 5; i++) {
        double acc = 0.;
#pragma unroll
        for (int j = 0; j < EVPGPU_NSYSMX; j++)
            if (j < ns) acc += in.sc[j][i] * r2[j];
        ed[i] = acc;
    }
    if (DD) {
#pragma unroll
        for (int i = 0; i < 5; i++)
#pragma unroll
            for (int j = i; j < 5; j++) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < EVPGPU_NSYSMX; k++)
                    if (k < ns) acc += in.sc[k][i] * in.sc[k][j] * r1[k];
                dd[i][j] = acc;
                dd[j][i] = acc;
            }
    }
}

// Crout LU with implicit pivoting (src/kinematics.c:282-354: same pivot
// choice, same 1e-5 singular-row and small-pivot rules) followed by one
// forward/back substitution for x = a^-1 b.  The reference forms the explicit
// inverse column by column (LU_inv_66) and multiplies it into the residual;
// its forward-substitution shortcut (src/kinematics.c:377-388) only skips
// exact zeros of the unit columns it inverts, so the plain substitution here
// is the same mathematics with 1/6 of the work.
// Returns false for a row of max |a| < 1e-5 (the reference exits with 35).
__device__ __forceinline__ bool lu_solve6(double a[6][6], double b[6])
{
    double vv[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        double big = 0.0;
#pragma unroll
        for (int j = 0; j < 6; j++) big = fabs(a[i][j]) > big ? fabs(a[i][j]) : big;
        if (fabs(big) < 1E-5) return false;
        vv[i] = 1.0 / big;
    }
#pragma unroll
    for (int j = 0; j < 6; j++) {
#pragma unroll
        for (int i = 0; i < j; i++) {
            double sum = a[i][j];
#pragma unroll
            for (int k = 0; k < i; k++) sum -= a[i][k] * a[k][j];
            a[i][j] = sum;
        }
        double big = 0.0;
        int imax = j;
#pragma unroll
        for (int i = j; i < 6; i++) {
            double sum = a[i][j];
#pragma unroll
            for (int k = 0; k < j; k++) sum -= a[i][k] * a[k][j];
            a[i][j] = sum;
            double dum = vv[i] * fabs(sum);
            if (dum >= big) {
                big = dum;
                imax = i;
            }
        }
        // row interchange j <-> imax, applied to the right-hand side as well
        // (equivalent to the reference's permutation during substitution)
#pragma unroll
        for (int r = j + 1; r < 6; r++) {
            if (r == imax) {
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    double d = a[r][k];
                    a[r][k] = a[j][k];
                    a[j][k] = d;
                }
                vv[r] = vv[j];
                double d = b[r];
                b[r] = b[j];
                b[j] = d;
            }
        }
        if (fabs(a[j][j]) < 1E-5) a[j][j] = 1.0E-20;
        if (j != 5) {
            double dum = 1.0 / (a[j][j]);
#pragma unroll
            for (int i = j + 1; i < 6; i++) a[i][j] *= dum;
        }
    }
#pragma unroll
    for (int i = 1; i < 6; i++) {
        double sum = b[i];
#pragma unroll
        for (int j = 0; j < i; j++) sum -= a[i][j] * b[j];
        b[i] = sum;
    }
#pragma unroll
    for (int i = 5; i >= 0; i--) {
        double sum = b[i];
#pragma unroll
        for (int j = i + 1; j < 6; j++) sum -= a[i][j] * b[j];
        b[i] = sum / a[i][i];
    }
    return true;
}

struct Vox {
    double sg[6][6];
    double xl[6], eps[6], str[6];
    double signorm; // Frobenius norm of lambda (src/evolution.c:92-96)
};

template <bool INTN>
__device__ __forceinline__ double nr_quality(const NewtonParams &P, int p, const RateIn &in,
                                             const Vox &V, const double sig[6],
                                             double gd[EVPGPU_NSYSMX])
{
    double ed[5], dd[5][5], tot[6], res[6];
    srate<false, INTN>(P, p, in, sig, ed, dd, gd);
#pragma unroll
    for (int i = 0; i < 6; i++) {
        tot[i] = V.eps[i] + (i < 5 ? ed[i] : 0.0) * P.dt;
#pragma unroll
        for (int j = 0; j < 6; j++) tot[i] += V.sg[i][j] * sig[j];
    }
#pragma unroll
    for (int i = 0; i < 6; i++) {
        res[i] = sig[i] - V.xl[i];
#pragma unroll
        for (int j = 0; j < 6; j++) res[i] += P.c066[i * 6 + j] * (tot[j] - V.str[j]);
    }
    double f = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++) f += res[i] * res[i];
    return f / 2.0;
}

// returns 0, EVPGPU_ERR_SINGULAR or EVPGPU_ERR_LINESEARCH
template <bool INTN>
__device__ int update_voxel(const NewtonParams &P, int p, const RateIn &in, const Vox &V,
                            double sig6[6], double edotp[6], double gd[EVPGPU_NSYSMX],
                            double &conv_s, double &conv_e, int &iterl, int &trials)
{
    trials = 0;
    double ed[5], dd[5][5], tot[6], res[6], old[6];
    conv_s = 0.0;
    conv_e = 0.0;
    iterl = 0;
    double erral = 10 * 1E-10;
#pragma unroll
    for (int i = 0; i < 6; i++) sig6[i] = V.xl[i];
    while ((iterl < 500) && (fabs(erral) > fabs(1E-10))) {
        iterl++;
#pragma unroll
        for (int i = 0; i < 6; i++) old[i] = sig6[i];
        srate<true, INTN>(P, p, in, sig6, ed, dd, gd);
#pragma unroll
        for (int i = 0; i < 6; i++) edotp[i] = i < 5 ? ed[i] : 0.0;
#pragma unroll
        for (int i = 0; i < 6; i++) {
            tot[i] = V.eps[i] + edotp[i] * P.dt;
#pragma unroll
            for (int j = 0; j < 6; j++) tot[i] += V.sg[i][j] * sig6[j];
        }
#pragma unroll
        for (int i = 0; i < 6; i++) {
            res[i] = sig6[i] - V.xl[i];
#pragma unroll
            for (int j = 0; j < 6; j++) res[i] += P.c066[i * 6 + j] * (tot[j] - V.str[j]);
        }
        double gradf[6], step[6];
        {
            double jac[6][6];
#pragma unroll
            for (int i = 0; i < 6; i++)
#pragma unroll
                for (int j = 0; j < 6; j++) {
                    double acc = (double)(i == j);
#pragma unroll
                    for (int k = 0; k < 6; k++) {
                        double m = (k < 5 && j < 5) ? (V.sg[k][j] + dd[k][j] * P.dt) : V.sg[k][j];
                        acc += P.c066[i * 6 + k] * m;
                    }
                    jac[i][j] = acc;
                }
#pragma unroll
            for (int i = 0; i < 6; i++) {
                double g = 0.0;
#pragma unroll
                for (int j = 0; j < 6; j++) g += res[j] * jac[j][i];
                gradf[i] = g;
                step[i] = res[i];
            }
            if (!lu_solve6(jac, step)) return EVPGPU_ERR_SINGULAR;
#pragma unroll
            for (int i = 0; i < 6; i++) step[i] = -step[i];
        }
        double f_old = 0.0;
#pragma unroll
        for (int i = 0; i < 6; i++) f_old += res[i] * res[i];
        f_old *= 0.5;
        const double stpmax = 10.0, ALF = 1.E-4, TOLX = 1.E-7;
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
        double alamin = TOLX / test, alam = 1.0, alam2 = 0.0, f2 = 0.0, tmplam = 0.0;
        while (true) {
#pragma unroll
            for (int i = 0; i < 6; i++) sig6[i] = old[i] + alam * step[i];
            double f_new = nr_quality<INTN>(P, p, in, V, sig6, gd);
            trials++;
            if (alam < alamin) {
#pragma unroll
                for (int i = 0; i < 6; i++) sig6[i] = old[i];
                break;
            } else if (f_new <= f_old + ALF * alam * slope) {
                break;
            } else {
                if (fabs(alam - 1.0) < 1.E-4) {
                    tmplam = -slope / (2.0 * (f_new - f_old - slope));
                } else {
                    double rhs1 = f_new - f_old - alam * slope;
                    double rhs2 = f2 - f_old - alam2 * slope;
                    double ac = (rhs1 / (alam * alam) - rhs2 / (alam2 * alam2)) / (alam - alam2);
                    double bc = (-alam2 * rhs1 / (alam * alam) + alam * rhs2 / (alam2 * alam2)) /
                                (alam - alam2);
                    if (fabs(ac) < 1E-4) {
                        tmplam = -slope / (2.0 * bc);
                    } else {
                        double disc = bc * bc - 3.0 * ac * slope;
                        if (disc < 0.0) tmplam = 0.5 * alam;
                        else if (bc <= 0.0) tmplam = (-bc + sqrt(disc)) / (3.0 * ac);
                        else tmplam = -slope / (bc + sqrt(disc));
                    }
                    if (tmplam > 0.5 * alam) tmplam = 0.5 * alam;
                }
            }
            alam2 = alam;
            f2 = f_new; // the backtrack value the cubic step needs (Q18)
            alam = tmplam > 0.1 * alam ? tmplam : 0.1 * alam;
        }
        double conv_nr = 0.0, ce = 0.0, cs = 0.0;
#pragma unroll
        for (int i = 0; i < 6; i++) {
            conv_nr += (sig6[i] - old[i]) * (sig6[i] - old[i]);
            ce += (tot[i] - V.str[i]) * (tot[i] - V.str[i]);
            cs += (sig6[i] - V.xl[i]) * (sig6[i] - V.xl[i]);
        }
        erral = sqrt(conv_nr) / V.signorm;
        conv_s = sqrt(cs);
        conv_e = sqrt(ce);
    }
    return 0;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

constexpr int NB_THREADS = 64;

// One lane per voxel.  Writes sig (in place), edot, gamdot; block partials
// (NRED doubles per block) of err_s, err_e, Newton iterations, sum sig,
// sum sig over phase 1.
template <bool INTN, bool CSYM>
__global__ __launch_bounds__(NB_THREADS) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_newton(
    NewtonParams P, long long N, const int *__restrict__ phase, const int *__restrict__ orient,
    const double *__restrict__ sgr, double *__restrict__ sig, const double *__restrict__ E,
    const double *__restrict__ epsp, double *__restrict__ edot, const double *__restrict__ schm,
    const double *__restrict__ crss, double *__restrict__ gamdot, double *__restrict__ partials,
    int *__restrict__ status, double *__restrict__ err_vox, int *__restrict__ niter_vox)
{
    long long v = (long long)blockIdx.x * NB_THREADS + threadIdx.x;
    double red[NRED];
#pragma unroll
    for (int i = 0; i < NRED; i++) red[i] = 0.0;
    if (v < N) {
        int p = phase[v] - 1;
        p = p < 0 ? 0 : (p > 1 ? 1 : p);
        if (!P.is_gas[p]) {
            Vox V;
            RateIn in;
            double t[6];
#pragma unroll
            for (int c = 0; c < 6; c++) t[c] = sig[c * N + v];
            // |lambda|_F over the full 3x3, row-major order
            V.signorm = sqrt(t[0] * t[0] + t[5] * t[5] + t[4] * t[4] + t[5] * t[5] + t[1] * t[1] +
                             t[3] * t[3] + t[4] * t[4] + t[3] * t[3] + t[2] * t[2]);
            sym_to_v6(t, V.xl);
#pragma unroll
            for (int c = 0; c < 6; c++) t[c] = epsp[c * N + v];
            sym_to_v6(t, V.eps);
#pragma unroll
            for (int c = 0; c < 6; c++) t[c] = E[c * N + v];
            sym_to_v6(t, V.str);
            const double *sgv = sgr + 36LL * orient[v];
#pragma unroll
            for (int i = 0; i < 6; i++)
#pragma unroll
                for (int j = 0; j < 6; j++) V.sg[i][j] = sgv[i * 6 + j];
            const int ns = P.nsys[p];
#pragma unroll
            for (int s = 0; s < EVPGPU_NSYSMX; s++) {
                if (s < ns) {
#pragma unroll
                    for (int j = 0; j < 5; j++) in.sc[s][j] = schm[(s * 5 + j) * N + v];
                    in.tf[s] = crss[(s * 2) * N + v];
                    in.tb[s] = CSYM ? in.tf[s] : crss[(s * 2 + 1) * N + v];
                } else {
#pragma unroll
                    for (int j = 0; j < 5; j++) in.sc[s][j] = 0.0;
                    in.tf[s] = in.tb[s] = 1.0;
                }
            }
            double s6[6], e6[6], gd[EVPGPU_NSYSMX], cs, ce;
            int it, ntr;
            int st = update_voxel<INTN>(P, p, in, V, s6, e6, gd, cs, ce, it, ntr);
            if (st) {
                atomicCAS(status, 0, st);
            } else {
                v6_to_sym(s6, t);
#pragma unroll
                for (int c = 0; c < 6; c++) sig[c * N + v] = t[c];
                v6_to_sym(e6, t);
#pragma unroll
                for (int c = 0; c < 6; c++) edot[c * N + v] = t[c];
#pragma unroll
                for (int s = 0; s < EVPGPU_NSYSMX; s++)
                    if (s < ns) gamdot[s * N + v] = gd[s];
                red[0] = cs;
                red[1] = ce;
                red[2] = (double)it;
                red[15] = (double)ntr;
                if (err_vox) {
                    err_vox[2 * v] = cs;
                    err_vox[2 * v + 1] = ce;
                }
                if (niter_vox) niter_vox[v] = it;
            }
        } else {
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
    __shared__ double wsum[NB_THREADS / 64][NRED];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NRED; i++) {
        double s = wave_sum(red[i]);
        if (lane == 0) wsum[w][i] = s;
    }
    __syncthreads();
    if (threadIdx.x < NRED) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < NB_THREADS / 64; q++) s += wsum[q][threadIdx.x];
        partials[(long long)blockIdx.x * NRED + threadIdx.x] = s;
    }
}

// Sum nblk rows of nval partials into out[nval], one thread per value,
// rows in order (deterministic).
__global__ void k_reduce_rows(const double *__restrict__ partials, int nblk, int nval,
                              double *__restrict__ out)
{
    __shared__ double sh[256];
    for (int val = 0; val < nval; val++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblk; b += 256) s += partials[(long long)b * nval + val];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[val] = sh[0];
        __syncthreads();
    }
}

} // namespace

int newton_launch(evpgpu_ctx *c, const NewtonParams &P, double *sig, const double *E,
                  const double *epsp, double *edot, const double *schm, const double *crss,
                  double *gamdot, const int *phase, const int *orient, const double *sgr,
                  long long n, double *partials, int *nblk_out, int *status, double *err_vox,
                  int *niter_vox)
{
    int nblk = (int)((n + NB_THREADS - 1) / NB_THREADS);
    // the integral-exponent variant needs every crystal phase to have one
    bool intn = true;
    for (int q = 0; q < 2; q++)
        if (!P.is_gas[q] && P.int_n[q] == 0) intn = false;
#define NEWTON_ARGS dim3(nblk), dim3(NB_THREADS), 0, c->stream, P, n, phase, orient, sgr, sig, E, \
                    epsp, edot, schm, crss, gamdot, partials, status, err_vox, niter_vox
    if (intn && P.crss_sym) hipLaunchKernelGGL((k_newton<true, true>), NEWTON_ARGS);
    else if (intn) hipLaunchKernelGGL((k_newton<true, false>), NEWTON_ARGS);
    else if (P.crss_sym) hipLaunchKernelGGL((k_newton<false, true>), NEWTON_ARGS);
    else hipLaunchKernelGGL((k_newton<false, false>), NEWTON_ARGS);
#undef NEWTON_ARGS
    if (nblk_out) *nblk_out = nblk;
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

int reduce_partials(evpgpu_ctx *c, const double *partials, int nblk, int nval, double *out)
{
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(256), 0, c->stream, partials, nblk, nval, out);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}
