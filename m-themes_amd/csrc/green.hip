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
#include "fft_core.h"
#include "green_point.h"
#include "spectral.h"

namespace {

using fftc::CT;
using kspace::unit_k;
using gop::gamma_apply;

// MODE 0: in-place symmetric output + antisymmetric accumulation (acc != null)
// MODE 1: write all 9 components to out9 (tests)
// MODE 2: mixed loop, delta spectrum: in place symmetric, acc += antisym + skA
// MODE 3: mixed loop, anchor spectrum: in place symmetric, skA = antisym
// The antisymmetric parts are accumulated at the split wave vectors only:
// everywhere else the increment's antisymmetric part follows from its
// symmetric part at the end of the increment (green_skew_fill).
template <typename R, int MODE>
__global__ __launch_bounds__(256) void k_green(typename CT<R>::T *spec, double2 *acc,
                                               double2 *out9, float2 *skA, GreenParams P, int nx,
                                               int ny, int nz, int nzh, long long NK)
{
    typedef typename CT<R>::T C;
    __shared__ R c0s[36];
    if (threadIdx.x < 36) c0s[threadIdx.x] = (R)P.ga[threadIdx.x];
    __syncthreads();
    long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= NK) return;
    int kz = (int)(id % nzh);
    int y = (int)((id / nzh) % ny);
    int x = (int)(id / ((long long)nzh * ny));
    C s6[6];
#pragma unroll
    for (int c = 0; c < 6; c++) s6[c] = spec[c * NK + id];
    if (MODE == 1) {
        C e[3][3];
        gop::green_tensor<R>(c0s, P.s0, P.elast_bc, x, y, kz, nx, ny, nz, s6, e);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                out9[(i * 3 + j) * NK + id] = make_double2((double)e[i][j].x, (double)e[i][j].y);
        return;
    }
    C e6[6];
    double wr[3], wi[3];
    bool split = false;
    if (id == 0) gop::green_zero<R>(c0s, P.s0, P.elast_bc, nx, ny, nz, s6, e6, wr, wi);
    else gop::green_point<R>(c0s, x, y, kz, nx, ny, nz, s6, e6, wr, wi, split);
#pragma unroll
    for (int c = 0; c < 6; c++) spec[c * NK + id] = e6[c];
    if (!split) return;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if (MODE == 3) {
            skA[c * NK + id] = make_float2((float)wr[c], (float)wi[c]);
        } else if (acc) {
            double2 w = acc[c * NK + id];
            double ax = wr[c], ay = wi[c];
            if (MODE == 2) {
                float2 k = skA[c * NK + id];
                ax += (double)k.x;
                ay += (double)k.y;
            }
            w.x += ax;
            w.y += ay;
            acc[c * NK + id] = w;
        }
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


// The antisymmetric part of an increment's Green outputs at a wave vector
// that is not split, from the spectrum of its symmetric part e (the
// increment's strain change): every output there is -u (x) n, so
// e = sym(a x n) with a = 2 e.n - (n.e.n) n and the antisymmetric part is
// skew(a x n) (23 13 12).  The split wave vectors and k = 0 keep the
// accumulated values.
__global__ __launch_bounds__(256) void k_skew_fill(const double2 *spec, double2 *acc, int nx, int ny, int nz,
                                                   int nzh, long long NK)
{
    long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= NK || id == 0) return;
    int kz = (int)(id % nzh);
    int y = (int)((id / nzh) % ny);
    int x = (int)(id / ((long long)nzh * ny));
    double n[3], n2[3];
    if (kspace::mode_dirs(x, y, kz, nx, ny, nz, n, n2)) return;
    double2 e[6];
    for (int c = 0; c < 6; c++) e[c] = spec[c * NK + id];
    // e.n with e symmetric (11 22 33 23 13 12)
    double2 en[3];
    en[0] = make_double2(e[0].x * n[0] + e[5].x * n[1] + e[4].x * n[2], e[0].y * n[0] + e[5].y * n[1] + e[4].y * n[2]);
    en[1] = make_double2(e[5].x * n[0] + e[1].x * n[1] + e[3].x * n[2], e[5].y * n[0] + e[1].y * n[1] + e[3].y * n[2]);
    en[2] = make_double2(e[4].x * n[0] + e[3].x * n[1] + e[2].x * n[2], e[4].y * n[0] + e[3].y * n[1] + e[2].y * n[2]);
    const double2 nen = make_double2(n[0] * en[0].x + n[1] * en[1].x + n[2] * en[2].x,
                                     n[0] * en[0].y + n[1] * en[1].y + n[2] * en[2].y);
    double2 a[3];
    for (int i = 0; i < 3; i++) a[i] = make_double2(2.0 * en[i].x - nen.x * n[i], 2.0 * en[i].y - nen.y * n[i]);
    acc[0 * NK + id] = make_double2(0.5 * (a[1].x * n[2] - a[2].x * n[1]), 0.5 * (a[1].y * n[2] - a[2].y * n[1]));
    acc[1 * NK + id] = make_double2(0.5 * (a[0].x * n[2] - a[2].x * n[0]), 0.5 * (a[0].y * n[2] - a[2].y * n[0]));
    acc[2 * NK + id] = make_double2(0.5 * (a[0].x * n[1] - a[1].x * n[0]), 0.5 * (a[0].y * n[1] - a[1].y * n[0]));
}
} // namespace

// acc at the wave vectors that are not split: the antisymmetric part of the
// increment from the spectrum of its strain change held in c->spec
int green_skew_fill(evpgpu_ctx *c)
{
    unsigned nb = (unsigned)((c->NK + 255) / 256);
    hipLaunchKernelGGL(k_skew_fill, dim3(nb), dim3(256), 0, c->stream, (const double2 *)c->spec, c->acc, c->nx,
                       c->ny, c->nz, c->nzh, c->NK);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

int green_project(evpgpu_ctx *c)
{
    unsigned nb = (unsigned)((c->NK + 255) / 256);
    hipLaunchKernelGGL(k_project, dim3(nb), dim3(256), 0, c->stream, (double2 *)c->spec, c->nx, c->ny,
                       c->nz, c->nzh, c->NK);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

GreenParams green_params(const evpgpu_ctx *c)
{
    GreenParams P;
    static const int pr[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};
    const double *C0 = c->ld.c0;
#define C4(i, j, k, l) C0[(((i)*3 + (j)) * 3 + (k)) * 3 + (l)]
    for (int q = 0; q < 6; q++)
        for (int p = 0; p < 6; p++) {
            const int i = pr[q][0], k = pr[q][1], j = pr[p][0], l = pr[p][1];
            P.ga[q * 6 + p] = j == l ? C4(i, j, k, l) : C4(i, j, k, l) + C4(i, l, k, j);
        }
#undef C4
    memcpy(P.s0, c->ld.s0, sizeof(P.s0));
    P.elast_bc = c->ld.elast_bc;
    return P;
}

// mixed loop: MODE 3 on the FP64 anchor spectrum, MODE 2 on the FP32 delta
int green_anchor(evpgpu_ctx *c)
{
    const GreenParams P = green_params(c);
    unsigned nb = (unsigned)((c->NK + 255) / 256);
    hipLaunchKernelGGL((k_green<double, 3>), dim3(nb), dim3(256), 0, c->stream, (double2 *)c->spec, nullptr,
                       nullptr, c->skA, P, c->nx, c->ny, c->nz, c->nzh, c->NK);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

int green_delta(evpgpu_ctx *c)
{
    const GreenParams P = green_params(c);
    unsigned nb = (unsigned)((c->NK + 255) / 256);
    hipLaunchKernelGGL((k_green<float, 2>), dim3(nb), dim3(256), 0, c->stream, (float2 *)c->spec, c->acc,
                       nullptr, c->skA, P, c->nx, c->ny, c->nz, c->nzh, c->NK);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

int green_apply(evpgpu_ctx *c, int precision, int accumulate)
{
    const GreenParams P = green_params(c);
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
    const GreenParams P = green_params(c);
    unsigned nb = (unsigned)((c->NK + 255) / 256);
    hipLaunchKernelGGL((k_green<double, 1>), dim3(nb), dim3(256), 0, c->stream, (double2 *)in6, nullptr,
                       out9, nullptr, P, c->nx, c->ny, c->nz, c->nzh, c->NK);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}
