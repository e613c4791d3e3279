// green_point.h -- the Green operator at one wave vector of the half
// spectrum: e-hat = Gamma_eff(k) : sigma-hat (src/evolution.c:1373-1387 with
// Gamma from src/init.c:541-591), shared by the Green pass (green.hip) and
// the fused x-line kernel (fft_fused.hip).
//
// Gamma is never stored: Gamma_ijkl = -Omega_ik n_j n_l with
// Omega = (C0_ijkl n_j n_l)^-1, so t = sigma-hat . n, u = Omega t,
// e-hat_ij = -u_i n_j.  Where a component of k sits at Nyquist, n(-k) is not
// -n(k) and Gamma_eff = (Gamma(n(k)) + Gamma(n(-k)))/2, which makes the C2R
// of the half spectrum equal the reference's Re(IFFT) of the full one.  At
// k = 0, -S0 : sigma-hat when ElastBC == 0, else zero.
#pragma once

#include <math.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define GPHD __host__ __device__ __forceinline__
#else
#define GPHD static inline
#endif

#include "spectral.h"

namespace gop {

// e = Gamma(n) s for complex s (3x3 full), e accumulated with weight w;
// ga: the 36 coefficients of the symmetric A = C0 : (n x n) (GreenParams)
template <typename R, typename C>
GPHD void gamma_apply(const R *ga, const double nd[3], const C s[3][3],
                                            C e[3][3], R w)
{
    R n[3] = {(R)nd[0], (R)nd[1], (R)nd[2]};
    const R nn[6] = {n[0] * n[0], n[1] * n[1], n[2] * n[2], n[1] * n[2], n[0] * n[2], n[0] * n[1]};
    R a[6];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        R acc = 0;
#pragma unroll
        for (int p = 0; p < 6; p++) acc += ga[q * 6 + p] * nn[p];
        a[q] = acc;
    }
    // M3Inverse (src/init.c:517-539) of the symmetric A: the adjugate O and
    // the determinant (A = [a0 a5 a4; a5 a1 a3; a4 a3 a2])
    R O[3][3];
    O[0][0] = a[1] * a[2] - a[3] * a[3];
    O[1][1] = a[0] * a[2] - a[4] * a[4];
    O[2][2] = a[0] * a[1] - a[5] * a[5];
    O[0][1] = O[1][0] = a[4] * a[3] - a[5] * a[2];
    O[0][2] = O[2][0] = a[5] * a[3] - a[4] * a[1];
    O[1][2] = O[2][1] = a[4] * a[5] - a[0] * a[3];
    R det = a[0] * O[0][0] + a[5] * O[0][1] + a[4] * O[0][2];
    R id = (R)1 / det;
    C t[3], u[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        t[k].x = s[k][0].x * n[0] + s[k][1].x * n[1] + s[k][2].x * n[2];
        t[k].y = s[k][0].y * n[0] + s[k][1].y * n[1] + s[k][2].y * n[2];
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        u[i].x = (O[i][0] * t[0].x + O[i][1] * t[1].x + O[i][2] * t[2].x) * id * w;
        u[i].y = (O[i][0] * t[0].y + O[i][1] * t[1].y + O[i][2] * t[2].y) * id * w;
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            e[i][j].x -= u[i].x * n[j];
            e[i][j].y -= u[i].y * n[j];
        }
}

// the full (non-symmetric) e-hat of one wave vector
template <typename R, typename C>
GPHD void green_tensor(const R *c0s, const double *s0, int elast_bc, int x, int y,
                                             int kz, int nx, int ny, int nz, const C s6[6], C e[3][3])
{
    C s[3][3] = {{s6[0], s6[5], s6[4]}, {s6[5], s6[1], s6[3]}, {s6[4], s6[3], s6[2]}};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) e[i][j] = C{(R)0, (R)0};
    if (x == 0 && y == 0 && kz == 0) {
        if (elast_bc == 0) {
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++)
#pragma unroll
                    for (int k = 0; k < 3; k++)
#pragma unroll
                        for (int l = 0; l < 3; l++) {
                            R g = -(R)s0[((i * 3 + j) * 3 + k) * 3 + l];
                            e[i][j].x += g * s[k][l].x;
                            e[i][j].y += g * s[k][l].y;
                        }
        }
        return;
    }
    double n1[3], n2[3];
    if (!kspace::mode_dirs(x, y, kz, nx, ny, nz, n1, n2)) {
        gamma_apply<R>(c0s, n1, s, e, (R)1);
    } else {
        gamma_apply<R>(c0s, n1, s, e, (R)0.5);
        gamma_apply<R>(c0s, n2, s, e, (R)0.5);
    }
}

// One direction's contribution to the symmetric (11 22 33 23 13 12) and
// antisymmetric (23 13 12) parts of e-hat = -w u (x) n, u = Omega (sigma-hat n),
// from the 6 stress components directly (no 3x3 tensors held: fewer registers
// in the fused x-line kernel)
template <typename R, typename C>
GPHD void gamma_acc6(const R *ga, const R nd[3], const C s6[6], C e6[6], C an[3],
                                           R w)
{
    const R n0 = nd[0], n1 = nd[1], n2 = nd[2];
    const R nn[6] = {n0 * n0, n1 * n1, n2 * n2, n1 * n2, n0 * n2, n0 * n1};
    R a[6];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        R acc = 0;
#pragma unroll
        for (int p = 0; p < 6; p++) acc += ga[q * 6 + p] * nn[p];
        a[q] = acc;
    }
    // symmetric adjugate and determinant of A = [a0 a5 a4; a5 a1 a3; a4 a3 a2]
    const R o00 = a[1] * a[2] - a[3] * a[3], o11 = a[0] * a[2] - a[4] * a[4], o22 = a[0] * a[1] - a[5] * a[5];
    const R o01 = a[4] * a[3] - a[5] * a[2], o02 = a[5] * a[3] - a[4] * a[1], o12 = a[4] * a[5] - a[0] * a[3];
    const R f = w / (a[0] * o00 + a[5] * o01 + a[4] * o02);
    // t = sigma-hat n (sigma symmetric: 11 22 33 23 13 12)
    C t0, t1, t2;
    t0.x = s6[0].x * n0 + s6[5].x * n1 + s6[4].x * n2;
    t0.y = s6[0].y * n0 + s6[5].y * n1 + s6[4].y * n2;
    t1.x = s6[5].x * n0 + s6[1].x * n1 + s6[3].x * n2;
    t1.y = s6[5].y * n0 + s6[1].y * n1 + s6[3].y * n2;
    t2.x = s6[4].x * n0 + s6[3].x * n1 + s6[2].x * n2;
    t2.y = s6[4].y * n0 + s6[3].y * n1 + s6[2].y * n2;
    C u0, u1, u2;
    u0.x = (o00 * t0.x + o01 * t1.x + o02 * t2.x) * f;
    u0.y = (o00 * t0.y + o01 * t1.y + o02 * t2.y) * f;
    u1.x = (o01 * t0.x + o11 * t1.x + o12 * t2.x) * f;
    u1.y = (o01 * t0.y + o11 * t1.y + o12 * t2.y) * f;
    u2.x = (o02 * t0.x + o12 * t1.x + o22 * t2.x) * f;
    u2.y = (o02 * t0.y + o12 * t1.y + o22 * t2.y) * f;
    const R h = (R)0.5;
    e6[0].x -= u0.x * n0; e6[0].y -= u0.y * n0;
    e6[1].x -= u1.x * n1; e6[1].y -= u1.y * n1;
    e6[2].x -= u2.x * n2; e6[2].y -= u2.y * n2;
    e6[3].x -= (u1.x * n2 + u2.x * n1) * h; e6[3].y -= (u1.y * n2 + u2.y * n1) * h;
    e6[4].x -= (u0.x * n2 + u2.x * n0) * h; e6[4].y -= (u0.y * n2 + u2.y * n0) * h;
    e6[5].x -= (u0.x * n1 + u1.x * n0) * h; e6[5].y -= (u0.y * n1 + u1.y * n0) * h;
    an[0].x -= (u1.x * n2 - u2.x * n1) * h; an[0].y -= (u1.y * n2 - u2.y * n1) * h;
    an[1].x -= (u0.x * n2 - u2.x * n0) * h; an[1].y -= (u0.y * n2 - u2.y * n0) * h;
    an[2].x -= (u0.x * n1 - u1.x * n0) * h; an[2].y -= (u0.y * n1 - u1.y * n0) * h;
}

// e6 / antisymmetric parts from a full e-hat
template <typename R, typename C>
GPHD void split_e(const C e[3][3], C e6[6], double wr[3], double wi[3])
{
    const int ia[6] = {0, 1, 2, 1, 0, 0}, ib[6] = {0, 1, 2, 2, 2, 1};
#pragma unroll
    for (int c = 0; c < 6; c++) {
        e6[c].x = (e[ia[c]][ib[c]].x + e[ib[c]][ia[c]].x) * (R)0.5;
        e6[c].y = (e[ia[c]][ib[c]].y + e[ib[c]][ia[c]].y) * (R)0.5;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        int a = ia[c + 3], b = ib[c + 3];
        wr[c] = (double)((e[a][b].x - e[b][a].x) * (R)0.5);
        wi[c] = (double)((e[a][b].y - e[b][a].y) * (R)0.5);
    }
}

// the k = 0 term: -S0 : sigma-hat when ElastBC == 0, else zero
template <typename R, typename C>
GPHD void green_zero(const R *ga, const double *s0, int elast_bc, int nx, int ny,
                                           int nz, const C s6[6], C e6[6], double wr[3], double wi[3])
{
    C e[3][3];
    green_tensor<R>(ga, s0, elast_bc, 0, 0, 0, nx, ny, nz, s6, e);
    split_e<R>(e, e6, wr, wi);
}

// symmetric part (11 22 33 23 13 12) and antisymmetric part (23 13 12, as
// FP64 real and imaginary parts) of e-hat at one wave vector k != 0 (the
// caller takes k = 0 to green_zero: kept apart, its code does not add to the
// registers of the per-point loops)
// split: the wave vector sits on a Nyquist plane where n(-k) != -n(k)
// (Gamma_eff mixes two directions there).  The directions are in R: FP64 the
// reference's arithmetic, FP32 the FP32 form of kspace::mode_dirs.
template <typename R, typename C>
GPHD void green_point(const R *ga, int x, int y, int kz, int nx, int ny, int nz,
                                            const C s6[6], C e6[6], double wr[3], double wi[3], bool &split)
{
    C an[3];
#pragma unroll
    for (int c = 0; c < 6; c++) e6[c] = C{(R)0, (R)0};
#pragma unroll
    for (int c = 0; c < 3; c++) an[c] = C{(R)0, (R)0};
    R n1[3], n2[3];
    split = kspace::mode_dirs(x, y, kz, nx, ny, nz, n1, n2);
    if (!split) {
        gamma_acc6<R>(ga, n1, s6, e6, an, (R)1);
    } else {
        gamma_acc6<R>(ga, n1, s6, e6, an, (R)0.5);
        gamma_acc6<R>(ga, n2, s6, e6, an, (R)0.5);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        wr[c] = (double)an[c].x;
        wi[c] = (double)an[c].y;
    }
}

} // namespace gop
