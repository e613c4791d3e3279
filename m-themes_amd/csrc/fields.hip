// fields.hip -- per-increment voxel kernels around the Green loop:
//   k_prologue  : DisGrad += Udot*dt (src/evolution.c:1318-1332)
//   k_epilogue  : Eps += Edot*dt, the Eps/Edot averages, update_orient and the
//                 Voce hardening, fused in one pass over the voxels
//                 (src/evolution.c:1469-1547, src/constitutive.c:142-326)
#include "evpgpu_impl.h"

namespace {

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ void k_fill(double *p, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ void k_fill_int(int *p, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_to_f32(const double *__restrict__ src, float *__restrict__ dst, long long n,
                         const int *__restrict__ skip)
{
    if (skip && *skip) return;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (float)src[i];
}

// plain streaming copy (16-byte accesses, grid-stride), the reference kernel
// for achievable HBM bandwidth and for scaling the traffic counters
__global__ __launch_bounds__(256) void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                              long long n2)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n2;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_sub(double *__restrict__ a, const double *__restrict__ b, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] -= b[i];
}

// b = a - b
__global__ void k_rsub(const double *__restrict__ a, double *__restrict__ b, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i] - b[i];
}

// E (6 comps) += dE (sym of Udot*dt), W (3 comps) += dW; creep: E = m6, W = w3
__global__ void k_prologue(double *E, double *W, long long N, double d0, double d1, double d2,
                           double d3, double d4, double d5, double w0, double w1, double w2,
                           int creep)
{
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    double d[6] = {d0, d1, d2, d3, d4, d5}, w[3] = {w0, w1, w2};
#pragma unroll
    for (int c = 0; c < 6; c++) E[c * N + v] = creep ? d[c] : E[c * N + v] + d[c];
#pragma unroll
    for (int c = 0; c < 3; c++) W[c * N + v] = creep ? w[c] : W[c * N + v] + w[c];
}

// dW (3 comps, fluctuation part from the inverse FFT, in rw) + uniform part m
// -> W += dW ; rw keeps dW for the epilogue
__global__ void k_antisym_add(double *W, double *rw, long long N, double m0, double m1, double m2)
{
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    double m[3] = {m0, m1, m2};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double d = m[c] + rw[c * N + v];
        rw[c * N + v] = d;
        W[c * N + v] += d;
    }
}

// Orientation (src/constitutive.c:142-194)
__device__ void orientation(double tran[3][3], double rot[3][3])
{
    double w[3] = {rot[2][1], rot[0][2], rot[1][0]};
    double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    double wt = tan(wn / 2.0);
    if (fabs(wn) < 1E-6) wn = 1.0;
#pragma unroll
    for (int i = 0; i < 3; i++) w[i] *= wt / wn;
    double om[3][3] = {{0., -1. * w[2], w[1]}, {w[2], 0., -1. * w[0]}, {-1. * w[1], w[0], 0.}};
    double om2[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < 3; i++) s += om[a][i] * om[i][b];
            om2[a][b] = s;
        }
    wt = wt * wt;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) rot[a][b] = (double)(a == b) + 2. * (om[a][b] + om2[a][b]) / (1. + wt);
    double tn[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < 3; i++) s += rot[a][i] * tran[i][b];
            tn[a][b] = s;
        }
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) tran[a][b] = tn[a][b];
}

__global__ __launch_bounds__(256) void k_epilogue(
    EpilogueParams P, long long N, const int *__restrict__ phase, double *__restrict__ epsp,
    const double *__restrict__ edot, const double *__restrict__ dW, double *__restrict__ tran,
    const double *__restrict__ gamdot, double *__restrict__ crss, double *__restrict__ gamacum,
    double *__restrict__ partials)
{
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double red[NRED_EPI];
#pragma unroll
    for (int i = 0; i < NRED_EPI; i++) red[i] = 0.0;
    if (v < N) {
        int p = phase[v] - 1;
        p = p < 0 ? 0 : (p > 1 ? 1 : p);
#pragma unroll
        for (int c = 0; c < 6; c++) {
            double e = epsp[c * N + v] + edot[c * N + v] * P.dt;
            epsp[c * N + v] = e;
            red[c] = e * P.wgt;
            red[6 + c] = edot[c * N + v] * P.wgt;
        }
        if (!P.is_gas[p]) {
            const int ns = P.nsys[p];
            double gd[EVPGPU_NSYSMX];
#pragma unroll
            for (int s = 0; s < EVPGPU_NSYSMX; s++) gd[s] = s < ns ? gamdot[s * N + v] : 0.0;
            if (P.update_orient) {
                double T[3][3], lp[3][3] = {{0}}, rl[3][3], rot[3][3];
#pragma unroll
                for (int i = 0; i < 9; i++) T[i / 3][i % 3] = tran[i * N + v];
                // lattice spin of the local velocity gradient: skew(L) = dW/dt
                double w23 = dW[0 * N + v] / P.dt, w13 = dW[1 * N + v] / P.dt,
                       w12 = dW[2 * N + v] / P.dt;
                rl[0][0] = rl[1][1] = rl[2][2] = 0.0;
                rl[0][1] = w12; rl[1][0] = -w12;
                rl[0][2] = w13; rl[2][0] = -w13;
                rl[1][2] = w23; rl[2][1] = -w23;
                for (int s = 0; s < ns; s++) {
                    double bs[3], nsv[3];
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        double a = 0.0, b = 0.0;
#pragma unroll
                        for (int j = 0; j < 3; j++) {
                            a += T[i][j] * P.bb[p][s][j];
                            b += T[i][j] * P.bn[p][s][j];
                        }
                        bs[i] = a;
                        nsv[i] = b;
                    }
#pragma unroll
                    for (int a = 0; a < 3; a++)
#pragma unroll
                        for (int b = 0; b < 3; b++) lp[a][b] += bs[a] * nsv[b] * gd[s];
                }
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int b = 0; b < 3; b++) {
                        double rs = (lp[a][b] - lp[b][a]) / 2.;
                        rot[a][b] = (P.udot_a[a * 3 + b] + rl[a][b] - rs) * P.dt;
                    }
                orientation(T, rot);
#pragma unroll
                for (int i = 0; i < 9; i++) tran[i * N + v] = T[i / 3][i % 3];
            }
            if (P.harden) {
                // harden() (src/constitutive.c:275-326); zero slip leaves the CRSS unchanged
                double gamtot = gamacum[v], deltgam = 0.0;
                for (int i = 0; i < ns; i++) deltgam += fabs(gd[i]) * P.dt;
                for (int i = 0; i < ns; i++) {
                    double dtau = 0.;
                    for (int j = 0; j < ns; j++) dtau += P.hard[p][i][j] * fabs(gd[j]) * P.dt;
                    double tau0 = P.tau[p][i][0], tau1 = P.tau[p][i][2];
                    double th0 = P.thet[p][i][0], th1 = P.thet[p][i][1];
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
                        double inc = dtau * voce / deltgam;
                        crss[(i * 2) * N + v] += inc;
                        crss[(i * 2 + 1) * N + v] += inc;
                    }
                }
                gamacum[v] = gamtot + deltgam;
            }
        }
    }
    __shared__ double ws[4][NRED_EPI];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NRED_EPI; i++) {
        double s = wsum(red[i]);
        if (lane == 0) ws[w][i] = s;
    }
    __syncthreads();
    if (threadIdx.x < NRED_EPI) {
        double s = ws[0][threadIdx.x] + ws[1][threadIdx.x] + ws[2][threadIdx.x] + ws[3][threadIdx.x];
        partials[(long long)blockIdx.x * NRED_EPI + threadIdx.x] = s;
    }
}

} // namespace

int fill_zero(evpgpu_ctx *c, double *p, long long n)
{
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, p, n);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

int fill_zero_int(evpgpu_ctx *c, int *p, long long n)
{
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, p, n);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

int to_f32_launch(evpgpu_ctx *c, const double *src, float *dst, long long n)
{
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, src, dst, n,
                       c->skip);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

int copy_launch(evpgpu_ctx *c, const double *src, double *dst, long long n)
{
    const long long n2 = n / 2;
    hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, c->stream, (const double2 *)src, (double2 *)dst, n2);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

int rsub_launch(evpgpu_ctx *c, const double *a, double *b, long long n)
{
    hipLaunchKernelGGL(k_rsub, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a, b, n);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

int sub_launch(evpgpu_ctx *c, double *a, const double *b, long long n)
{
    hipLaunchKernelGGL(k_sub, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a, b, n);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

int prologue_launch(evpgpu_ctx *c, const double *dE6, int creep, const double *macro_sym6)
{
    (void)macro_sym6;
    const double *w = dE6 + 6;
    hipLaunchKernelGGL(k_prologue, dim3((unsigned)((c->N + 255) / 256)), dim3(256), 0, c->stream, c->E,
                       c->W, c->N, dE6[0], dE6[1], dE6[2], dE6[3], dE6[4], dE6[5], w[0], w[1], w[2],
                       creep);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

int antisym_add_launch(evpgpu_ctx *c, const double *m3)
{
    hipLaunchKernelGGL(k_antisym_add, dim3((unsigned)((c->N + 255) / 256)), dim3(256), 0, c->stream, c->W,
                       c->rwork, c->N, m3[0], m3[1], m3[2]);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

int epilogue_launch(evpgpu_ctx *c, const EpilogueParams &P, double *partials, int *nblk)
{
    int nb = (int)((c->N + 255) / 256);
    hipLaunchKernelGGL(k_epilogue, dim3(nb), dim3(256), 0, c->stream, P, c->N, c->phase, c->epsp, c->edot,
                       c->rwork, c->tran, c->gamdot, c->crss, c->gamacum, partials);
    *nblk = nb;
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}
