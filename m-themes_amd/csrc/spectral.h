// spectral.h -- wave vectors of the half spectrum, shared by the Green kernel
// and the adjoint's spectral operators.
#pragma once

#include <math.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define KSHD __host__ __device__ __forceinline__
#else
#define KSHD static inline
#endif

namespace kspace {

// unit vector of xi = 2 pi (k/N - round(k/N)) as src/init.c:775-796 builds it
// (the Nyquist index maps to -pi; k = 0 gives the zero vector)
KSHD void unit_k(int x, int y, int z, int nx, int ny, int nz, double n[3])
{
    const double TWO_PI_REF = 6.28318530718; // src/evp.h:78
    double t[3] = {(double)x * (double)(1.0 / nx), (double)y * (double)(1.0 / ny),
                   (double)z * (double)(1.0 / nz)};
#pragma unroll
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
    n[0] = t[0] * s;
    n[1] = t[1] * s;
    n[2] = t[2] * s;
}

// The directions of half-spectrum entry (x, y, kz): n1 = n(k), and at
// Nyquist entries n2 = n(-k).  Returns true where n2 != +-n1, i.e. where the
// Hermitian-symmetrised Green operator mixes two directions.
KSHD bool mode_dirs(int x, int y, int kz, int nx, int ny, int nz, double n1[3], double n2[3])
{
    unit_k(x, y, kz, nx, ny, nz, n1);
    n2[0] = n1[0];
    n2[1] = n1[1];
    n2[2] = n1[2];
    if ((2 * x == nx) || (2 * y == ny) || (2 * kz == nz)) {
        unit_k((nx - x) % nx, (ny - y) % ny, (nz - kz) % nz, nx, ny, nz, n2);
        bool same = (n2[0] == n1[0] && n2[1] == n1[1] && n2[2] == n1[2]) ||
                    (n2[0] == -n1[0] && n2[1] == -n1[1] && n2[2] == -n1[2]);
        return !same;
    }
    return false;
}

// The same directions in FP32 for the FP32 Green pass.  xi_m is the signed
// index (k_m, less N_m from the Nyquist index on) times 1/N_m: the common
// factor 2 pi cancels in n = xi/|xi|, and n is within 2e-7 of the FP64
// direction.  The split test needs no directions: n(-k) keeps the Nyquist
// components of n(k) and negates the others, so n(-k) = +-n(k) exactly when
// no component is Nyquist or every non-Nyquist one is zero.
KSHD bool mode_dirs(int x, int y, int kz, int nx, int ny, int nz, float n1[3], float n2[3])
{
    const int id[3] = {x, y, kz}, len[3] = {nx, ny, nz};
    float t[3];
    bool nyq[3], any_nyq = false, other = false;
#pragma unroll
    for (int m = 0; m < 3; m++) {
        const int s = 2 * id[m] >= len[m] ? id[m] - len[m] : id[m];
        t[m] = (float)s * (1.f / (float)len[m]);
        nyq[m] = 2 * id[m] == len[m];
        any_nyq |= nyq[m];
        other |= id[m] != 0 && !nyq[m];
    }
    const float g2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
#ifdef __HIP_DEVICE_COMPILE__
    const float r = rsqrtf(g2);
#else
    const float r = 1.f / sqrtf(g2);
#endif
#pragma unroll
    for (int m = 0; m < 3; m++) {
        n1[m] = t[m] * r;
        n2[m] = nyq[m] ? n1[m] : -n1[m];
    }
    return any_nyq && other;
}

} // namespace kspace
