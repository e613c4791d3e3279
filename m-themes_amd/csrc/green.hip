// green.hip -- the Green convolution e-hat = Gamma-hat : sigma-hat on the half
// spectrum (src/evolution.c:1373-1387 with Gamma from src/init.c:541-591).
//
// Gamma is never stored.  It factorises as Gamma_ijkl = -Omega_ik n_j n_l with
// Omega = (C0_ijkl n_j n_l)^-1, so per wave vector
//     t = sigma-hat . n,   u = Omega t,   e-hat_ij = -u_i n_j,
// about 120 flops instead of the 81 complex MACs and 648 B/voxel of the
// reference's dense GAMMA.  n is the unit vector of
// xi = 2 pi (k/N - round(k/N)) as src/init.c:775-796 builds it (the Nyquist
// index maps to -pi).
//
// The reference transforms complex data and keeps Re(IFFT).  Where a
// component of k sits at Nyquist, n(-k) is not -n(k), so Gamma(k) and
// Gamma(-k) differ, and Re(IFFT(Gamma sigma-hat)) equals the Hermitian
// spectrum Gamma_eff sigma-hat with Gamma_eff = (Gamma(n(k)) + Gamma(n(-k)))/2.
// This kernel applies Gamma_eff, so the C2R result equals the reference's
// real part exactly in exact arithmetic.
//
// Output: the 6 symmetric components replace sigma-hat in place (11 22 33 23
// 13 12); the 3 antisymmetric components (23 13 12) are added to an FP64
// accumulator that the increment epilogue inverts once (they only feed the
// lattice spin of update_orient).
#include "evpgpu_impl.h"
#include "spectral.h"

namespace {

template <typename R> struct CT;
template <> struct CT<double> { typedef double2 T; };
template <> struct CT<float> { typedef float2 T; };

using kspace::unit_k;

// e = Gamma(n) s for complex s (3x3 full), e accumulated with weight w
template <typename R, typename C>
__device__ __forceinline__ void gamma_apply(const R *c0, const double nd[3], const C s[3][3],
                                            C e[3][3], R w)
{
    R n[3] = {(R)nd[0], (R)nd[1], (R)nd[2]};
    R A[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            R acc = 0;
#pragma unroll
            for (int j = 0; j < 3; j++)
#pragma unroll
                for (int l = 0; l < 3; l++) acc += c0[((i * 3 + j) * 3 + k) * 3 + l] * n[j] * n[l];
            A[i][k] = acc;
        }
    // M3Inverse (src/init.c:517-539)
    R det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
            A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
            A[0][2] * (A[1][0] * A[2][1] - A[2][0] * A[1][1]);
    R O[3][3];
    O[0][0] = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    O[0][1] = A[0][2] * A[2][1] - A[0][1] * A[2][2];
    O[0][2] = A[0][1] * A[1][2] - A[0][2] * A[1][1];
    O[1][0] = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    O[1][1] = A[0][0] * A[2][2] - A[0][2] * A[2][0];
    O[1][2] = A[0][2] * A[1][0] - A[0][0] * A[1][2];
    O[2][0] = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    O[2][1] = A[0][1] * A[2][0] - A[0][0] * A[2][1];
    O[2][2] = A[0][0] * A[1][1] - A[0][1] * A[1][0];
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

// MODE 0: in-place symmetric output + antisymmetric accumulation (acc != null)
// MODE 1: write all 9 components to out9 (tests)
// MODE 2: mixed loop, delta spectrum: in place symmetric, acc += antisym + skA
// MODE 3: mixed loop, anchor spectrum: in place symmetric, skA = antisym
template <typename R, int MODE>
__global__ __launch_bounds__(256) void k_green(typename CT<R>::T *spec, double2 *acc,
                                               double2 *out9, float2 *skA, GreenParams P, int nx,
                                               int ny, int nz, int nzh, long long NK)
{
    typedef typename CT<R>::T C;
    __shared__ R c0s[81];
    if (threadIdx.x < 81) c0s[threadIdx.x] = (R)P.c0[threadIdx.x];
    __syncthreads();
    long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= NK) return;
    int kz = (int)(id % nzh);
    int y = (int)((id / nzh) % ny);
    int x = (int)(id / ((long long)nzh * ny));
    C s6[6];
#pragma unroll
    for (int c = 0; c < 6; c++) s6[c] = spec[c * NK + id];
    C s[3][3] = {{s6[0], s6[5], s6[4]}, {s6[5], s6[1], s6[3]}, {s6[4], s6[3], s6[2]}};
    C e[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) e[i][j] = C{(R)0, (R)0};
    if (x == 0 && y == 0 && kz == 0) {
        if (P.elast_bc == 0) {
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++)
#pragma unroll
                    for (int k = 0; k < 3; k++)
#pragma unroll
                        for (int l = 0; l < 3; l++) {
                            R g = -(R)P.s0[((i * 3 + j) * 3 + k) * 3 + l];
                            e[i][j].x += g * s[k][l].x;
                            e[i][j].y += g * s[k][l].y;
                        }
        }
    } else {
        double n1[3];
        unit_k(x, y, kz, nx, ny, nz, n1);
        bool nyq = (2 * x == nx) || (2 * y == ny) || (2 * kz == nz);
        bool split = false;
        double n2[3];
        if (nyq) {
            unit_k((nx - x) % nx, (ny - y) % ny, (nz - kz) % nz, nx, ny, nz, n2);
            bool same = (n2[0] == n1[0] && n2[1] == n1[1] && n2[2] == n1[2]) ||
                        (n2[0] == -n1[0] && n2[1] == -n1[1] && n2[2] == -n1[2]);
            split = !same;
        }
        if (!split) {
            gamma_apply<R>(c0s, n1, s, e, (R)1);
        } else {
            gamma_apply<R>(c0s, n1, s, e, (R)0.5);
            gamma_apply<R>(c0s, n2, s, e, (R)0.5);
        }
    }
    if (MODE != 1) {
        const int ia[6] = {0, 1, 2, 1, 0, 0}, ib[6] = {0, 1, 2, 2, 2, 1};
#pragma unroll
        for (int c = 0; c < 6; c++) {
            C v;
            v.x = (e[ia[c]][ib[c]].x + e[ib[c]][ia[c]].x) * (R)0.5;
            v.y = (e[ia[c]][ib[c]].y + e[ib[c]][ia[c]].y) * (R)0.5;
            spec[c * NK + id] = v;
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            int a = ia[c + 3], b = ib[c + 3];
            double wx = (double)((e[a][b].x - e[b][a].x) * (R)0.5);
            double wy = (double)((e[a][b].y - e[b][a].y) * (R)0.5);
            if (MODE == 3) {
                skA[c * NK + id] = make_float2((float)wx, (float)wy);
            } else if (acc) {
                double2 w = acc[c * NK + id];
                if (MODE == 2) {
                    float2 k = skA[c * NK + id];
                    wx += (double)k.x;
                    wy += (double)k.y;
                }
                w.x += wx;
                w.y += wy;
                acc[c * NK + id] = w;
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                out9[(i * 3 + j) * NK + id] = make_double2((double)e[i][j].x, (double)e[i][j].y);
    }
}

// Orthogonal projection of a symmetric strain spectrum onto compatible
// fields sym(a x n): a = 2 e.n - (n.e.n) n (exact for e = sym(a x n)).  The
// k = 0 mode (macroscopic strain) is left unchanged.
__global__ __launch_bounds__(256) void k_project(double2 *spec, int nx, int ny, int nz, int nzh,
                                                 long long NK)
{
    long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= NK) return;
    int kz = (int)(id % nzh);
    int y = (int)((id / nzh) % ny);
    int x = (int)(id / ((long long)nzh * ny));
    if (x == 0 && y == 0 && kz == 0) return;
    double n[3], n2[3];
    unit_k(x, y, kz, nx, ny, nz, n);
    if (kspace::mode_dirs(x, y, kz, nx, ny, nz, n, n2)) {
        // Gamma_eff mixes n and n2 here: the compatible set is
        // {sym(a x n + b x n2)}, whose orthogonal complement is m x m with
        // m = n x n2, so remove that one component
        double m[3] = {n[1] * n2[2] - n[2] * n2[1], n[2] * n2[0] - n[0] * n2[2],
                       n[0] * n2[1] - n[1] * n2[0]};
        double mm = 1.0 / sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        for (int i = 0; i < 3; i++) m[i] *= mm;
        const double w[6] = {m[0] * m[0], m[1] * m[1], m[2] * m[2], m[1] * m[2], m[0] * m[2], m[0] * m[1]};
        double2 v[6], s = make_double2(0.0, 0.0);
        for (int c = 0; c < 6; c++) {
            v[c] = spec[c * NK + id];
            const double f = c < 3 ? 1.0 : 2.0; // m.e.m: off-diagonal entries count twice
            s.x += f * w[c] * v[c].x;
            s.y += f * w[c] * v[c].y;
        }
        for (int c = 0; c < 6; c++)
            spec[c * NK + id] = make_double2(v[c].x - s.x * w[c], v[c].y - s.y * w[c]);
        return;
    }
    double2 v[6];
#pragma unroll
    for (int c = 0; c < 6; c++) v[c] = spec[c * NK + id];
    double2 e[3][3] = {{v[0], v[5], v[4]}, {v[5], v[1], v[3]}, {v[4], v[3], v[2]}};
    double2 en[3], a[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        en[i].x = e[i][0].x * n[0] + e[i][1].x * n[1] + e[i][2].x * n[2];
        en[i].y = e[i][0].y * n[0] + e[i][1].y * n[1] + e[i][2].y * n[2];
    }
    double2 nen;
    nen.x = en[0].x * n[0] + en[1].x * n[1] + en[2].x * n[2];
    nen.y = en[0].y * n[0] + en[1].y * n[1] + en[2].y * n[2];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        a[i].x = 2.0 * en[i].x - nen.x * n[i];
        a[i].y = 2.0 * en[i].y - nen.y * n[i];
    }
    const int ia[6] = {0, 1, 2, 1, 0, 0}, ib[6] = {0, 1, 2, 2, 2, 1};
#pragma unroll
    for (int c = 0; c < 6; c++) {
        int i = ia[c], j = ib[c];
        spec[c * NK + id] = make_double2(0.5 * (a[i].x * n[j] + a[j].x * n[i]),
                                         0.5 * (a[i].y * n[j] + a[j].y * n[i]));
    }
}

} // namespace

int green_project(evpgpu_ctx *c)
{
    unsigned nb = (unsigned)((c->NK + 255) / 256);
    hipLaunchKernelGGL(k_project, dim3(nb), dim3(256), 0, c->stream, (double2 *)c->spec, c->nx, c->ny,
                       c->nz, c->nzh, c->NK);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

// mixed loop: MODE 3 on the FP64 anchor spectrum, MODE 2 on the FP32 delta
int green_anchor(evpgpu_ctx *c)
{
    GreenParams P;
    memcpy(P.c0, c->ld.c0, sizeof(P.c0));
    memcpy(P.s0, c->ld.s0, sizeof(P.s0));
    P.elast_bc = c->ld.elast_bc;
    unsigned nb = (unsigned)((c->NK + 255) / 256);
    hipLaunchKernelGGL((k_green<double, 3>), dim3(nb), dim3(256), 0, c->stream, (double2 *)c->spec, nullptr,
                       nullptr, c->skA, P, c->nx, c->ny, c->nz, c->nzh, c->NK);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

int green_delta(evpgpu_ctx *c)
{
    GreenParams P;
    memcpy(P.c0, c->ld.c0, sizeof(P.c0));
    memcpy(P.s0, c->ld.s0, sizeof(P.s0));
    P.elast_bc = c->ld.elast_bc;
    unsigned nb = (unsigned)((c->NK + 255) / 256);
    hipLaunchKernelGGL((k_green<float, 2>), dim3(nb), dim3(256), 0, c->stream, (float2 *)c->spec, c->acc,
                       nullptr, c->skA, P, c->nx, c->ny, c->nz, c->nzh, c->NK);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

int green_apply(evpgpu_ctx *c, int precision, int accumulate)
{
    GreenParams P;
    memcpy(P.c0, c->ld.c0, sizeof(P.c0));
    memcpy(P.s0, c->ld.s0, sizeof(P.s0));
    P.elast_bc = c->ld.elast_bc;
    unsigned nb = (unsigned)((c->NK + 255) / 256);
    if (precision == EVPGPU_MIXED)
        hipLaunchKernelGGL((k_green<float, 0>), dim3(nb), dim3(256), 0, c->stream, (float2 *)c->spec,
                           accumulate ? c->acc : nullptr, nullptr, nullptr, P, c->nx, c->ny, c->nz, c->nzh,
                           c->NK);
    else
        hipLaunchKernelGGL((k_green<double, 0>), dim3(nb), dim3(256), 0, c->stream, (double2 *)c->spec,
                           accumulate ? c->acc : nullptr, nullptr, nullptr, P, c->nx, c->ny, c->nz, c->nzh,
                           c->NK);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

int green_full9(evpgpu_ctx *c, const double2 *in6, double2 *out9)
{
    GreenParams P;
    memcpy(P.c0, c->ld.c0, sizeof(P.c0));
    memcpy(P.s0, c->ld.s0, sizeof(P.s0));
    P.elast_bc = c->ld.elast_bc;
    unsigned nb = (unsigned)((c->NK + 255) / 256);
    hipLaunchKernelGGL((k_green<double, 1>), dim3(nb), dim3(256), 0, c->stream, (double2 *)in6, nullptr,
                       out9, nullptr, P, c->nx, c->ny, c->nz, c->nzh, c->NK);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}
