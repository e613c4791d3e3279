// drx.hip -- the DRX nucleation check and the static nucleation strengths of
// M-THEMES' recrystallisation coupling, on one GPU.
//
//   evpgpu_drx_nucleate: NucleationCheckDRX (src/integrated.c:161-494) in the
//     branch the shipped build compiles (GB_BULGING_ONLY, Nucl_Static_Flag 1),
//     first call: per voxel the total dislocation density, the grain-boundary
//     indicator and the test against the static strength (one kernel), the
//     nuclei numbered in sweep order (a scan), and the spheres of the new
//     grains painted (the last nucleus in sweep order within the radius wins,
//     as the reference's loop over the nucleus list leaves it).
//   evpgpu_drx_weibull: the nucleation strengths kappa_drx =
//     gsl_ran_weibull(r, k_c, alpha) drawn voxel by voxel from GSL's
//     MT19937 (src/init.c:1008-1051) with a given seed -- a sequential 32-bit
//     stream, drawn on the host.
//
// Built with -ffp-contract=off: the density sums are the reference's
// separately rounded products and sums.
#include <hipcub/hipcub.hpp>
#include <math.h>

#include "evpgpu_impl.h"

namespace {

constexpr int NT = 256;

int grid_for(long long n)
{
    long long b = (n + NT - 1) / NT;
    if (b > 65536) b = 65536;
    return (int)(b < 1 ? 1 : b);
}

struct NucIn {
    int Lx, Ly, Lz, nsys;
    long long N;
    double kc_new; // zeta_kc * k_c
};

// src/integrated.c:181-193 (total density, NUCL_RHO_DIFF off) and 264-300
// (the test), with GB_indicator of src/evolution.c:816-828 from the checks
// of src/constitutive.c:790-853 (the sixth compares with -y)
__global__ void k_nuc(NucIn P, const int *grain, const double *rs, const double *rm, const double *g1,
                      const double *g2, const double *g3, const double *kappa, int *flag, int *gid_rex,
                      int *nuclei, double *kc, double *rho_tot)
{
    const long long N = P.N;
    for (long long v = blockIdx.x * (long long)NT + threadIdx.x; v < N; v += (long long)gridDim.x * NT) {
        double r = 0.0;
        for (int i = 0; i < P.nsys; i++) {
            const long long q = (long long)i * N + v;
            r += rs[q] + rm[q] + sqrt(g1[q] * g1[q] + g2[q] * g2[q] + g3[q] * g3[q]);
        }
        rho_tot[v] = r;
        const double diff_rho = r / 12;
        const int z = (int)(v % P.Lz);
        const long long t = v / P.Lz;
        const int y = (int)(t % P.Ly), x = (int)(t / P.Ly);
        const int xp = (x + 1) % P.Lx, xm = (x + P.Lx - 1) % P.Lx, yp = (y + 1) % P.Ly, ym = (y + P.Ly - 1) % P.Ly;
        const int zp = (z + 1) % P.Lz;
        auto at = [&](int a, int b, int c) { return grain[((long long)a * P.Ly + b) * P.Lz + c]; };
        const int g = grain[v];
        const int chk = (g == at(xp, y, z)) + (g == at(xm, y, z)) + (g == at(x, yp, z)) + (g == at(x, ym, z)) +
                        (g == at(x, y, zp)) + (g == at(x, ym, z));
        const int f = chk < 6 && diff_rho > kappa[v];
        flag[v] = f;
        nuclei[v] = f;
        if (f) {
            gid_rex[v] = 1;
            kc[v] = P.kc_new;
        }
    }
}

// src/integrated.c:318-405: nucleus k in sweep order gets grain n_grains + 1 + k
// and paints |r - r_k|^2 <= radius (no periodic wrap); the last painter wins
__global__ void k_paint(NucIn P, const int *flag, const long long *rank, int n_grains, int radius, int h,
                        const int *grain_in, int *grain_out, int *gid_rex, int *nuclei, long long *vox,
                        long long max_new)
{
    const long long N = P.N;
    for (long long u = blockIdx.x * (long long)NT + threadIdx.x; u < N; u += (long long)gridDim.x * NT) {
        const int z = (int)(u % P.Lz);
        const long long t = u / P.Lz;
        const int y = (int)(t % P.Ly), x = (int)(t / P.Ly);
        long long best = -1;
        for (int a = x - h; a <= x + h; a++)
            for (int b = y - h; b <= y + h; b++)
                for (int c = z - h; c <= z + h; c++) {
                    if (a < 0 || a >= P.Lx || b < 0 || b >= P.Ly || c < 0 || c >= P.Lz) continue;
                    if ((a - x) * (a - x) + (b - y) * (b - y) + (c - z) * (c - z) > radius) continue;
                    const long long v = ((long long)a * P.Ly + b) * P.Lz + c;
                    if (flag[v] && v > best) best = v;
                }
        if (best >= 0) {
            grain_out[u] = n_grains + 1 + (int)rank[best];
            gid_rex[u] = 1;
            nuclei[u] = 1;
        } else {
            grain_out[u] = grain_in[u];
        }
        if (flag[u] && rank[u] < max_new) vox[rank[u]] = u;
    }
}

// GSL's mt19937 (Matsumoto and Nishimura's generator with the 2002 seeding,
// seed 0 taken as 4357) and gsl_rng_uniform_pos / gsl_ran_weibull
struct MT {
    uint32_t s[624];
    int i;
    explicit MT(unsigned long seed)
    {
        if (seed == 0) seed = 4357;
        s[0] = (uint32_t)(seed & 0xffffffffUL);
        for (int k = 1; k < 624; k++) s[k] = (uint32_t)(1812433253UL * (s[k - 1] ^ (s[k - 1] >> 30)) + k);
        i = 624;
    }
    uint32_t next()
    {
        if (i >= 624) {
            for (int k = 0; k < 624; k++) {
                const uint32_t y = (s[k] & 0x80000000u) | (s[(k + 1) % 624] & 0x7fffffffu);
                s[k] = s[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
            }
            i = 0;
        }
        uint32_t k = s[i++];
        k ^= k >> 11;
        k ^= (k << 7) & 0x9d2c5680u;
        k ^= (k << 15) & 0xefc60000u;
        k ^= k >> 18;
        return k;
    }
    double uniform_pos()
    {
        double x;
        do x = next() / 4294967296.0;
        while (x == 0);
        return x;
    }
};

} // namespace

extern "C" int evpgpu_drx_nucleate(int Lx, int Ly, int Lz, int nsys, const int *grain_in, const double *rho_s,
                                   const double *rho_m, const double *g1, const double *g2, const double *g3,
                                   const double *kappa, int n_grains, int nucleus_radius, double zeta_kc,
                                   double k_c, int *grain_out, int *gid_rex, int *nuclei, double *kc,
                                   double *rho_tot, long long *nucleus_vox, int max_new, int *n_new)
{
    if (Lx < 1 || Ly < 1 || Lz < 1 || nsys < 0 || nsys > EVPGPU_NSYSMX || !grain_in || !kappa || !grain_out ||
        !gid_rex || !nuclei || !kc || !rho_tot || !n_new || max_new < 0 || (max_new > 0 && !nucleus_vox) ||
        nucleus_radius < 0 || (nsys > 0 && (!rho_s || !rho_m || !g1 || !g2 || !g3)))
        return EVPGPU_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return EVPGPU_ERR_NODEV;
    NucIn P;
    P.Lx = Lx;
    P.Ly = Ly;
    P.Lz = Lz;
    P.nsys = nsys;
    P.N = (long long)Lx * Ly * Lz;
    P.kc_new = zeta_kc * k_c;
    const long long N = P.N, M = (long long)nsys * N;
    int *d_grain = nullptr, *d_out = nullptr, *d_flag = nullptr, *d_rex = nullptr, *d_nuc = nullptr;
    double *d_r[5] = {nullptr, nullptr, nullptr, nullptr, nullptr}, *d_kappa = nullptr, *d_kc = nullptr,
           *d_rho = nullptr;
    long long *d_rank = nullptr, *d_vox = nullptr;
    void *tmp = nullptr;
    hipStream_t s = nullptr;
    int st = EVPGPU_OK;
    auto A = [&](void **p, size_t b) {
        if (st == EVPGPU_OK && hipMalloc(p, b > 0 ? b : 1) != hipSuccess) st = EVPGPU_ERR_NOMEM;
    };
    A((void **)&d_grain, N * sizeof(int));
    A((void **)&d_out, N * sizeof(int));
    A((void **)&d_flag, N * sizeof(int));
    A((void **)&d_rex, N * sizeof(int));
    A((void **)&d_nuc, N * sizeof(int));
    for (int k = 0; k < 5; k++) A((void **)&d_r[k], M * sizeof(double));
    A((void **)&d_kappa, N * sizeof(double));
    A((void **)&d_kc, N * sizeof(double));
    A((void **)&d_rho, N * sizeof(double));
    A((void **)&d_rank, N * sizeof(long long));
    A((void **)&d_vox, (max_new > 0 ? max_new : 1) * sizeof(long long));
    if (st == EVPGPU_OK && hipStreamCreate(&s) != hipSuccess) st = EVPGPU_ERR_HIP;
    const double *hr[5] = {rho_s, rho_m, g1, g2, g3};
    auto H2D = [&](void *d, const void *h, size_t b) {
        if (st == EVPGPU_OK && b && hipMemcpyAsync(d, h, b, hipMemcpyHostToDevice, s) != hipSuccess)
            st = EVPGPU_ERR_HIP;
    };
    H2D(d_grain, grain_in, N * sizeof(int));
    for (int k = 0; k < 5; k++) H2D(d_r[k], hr[k], M * sizeof(double));
    H2D(d_kappa, kappa, N * sizeof(double));
    H2D(d_kc, kc, N * sizeof(double));
    H2D(d_rex, gid_rex, N * sizeof(int));
    size_t need = 0;
    if (st == EVPGPU_OK) {
        hipLaunchKernelGGL(k_nuc, dim3(grid_for(N)), dim3(NT), 0, s, P, d_grain, d_r[0], d_r[1], d_r[2], d_r[3],
                           d_r[4], d_kappa, d_flag, d_rex, d_nuc, d_kc, d_rho);
        if (hipcub::DeviceScan::ExclusiveSum(nullptr, need, d_flag, d_rank, N, s) != hipSuccess) st = EVPGPU_ERR_HIP;
    }
    A(&tmp, need);
    if (st == EVPGPU_OK &&
        hipcub::DeviceScan::ExclusiveSum(tmp, need, d_flag, d_rank, N, s) != hipSuccess)
        st = EVPGPU_ERR_HIP;
    long long total = 0;
    if (st == EVPGPU_OK) {
        long long last = 0;
        int lastf = 0;
        if (hipMemcpyAsync(&last, d_rank + N - 1, sizeof(long long), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(&lastf, d_flag + N - 1, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            st = EVPGPU_ERR_HIP;
        total = last + lastf;
    }
    if (st == EVPGPU_OK) {
        int h = 0;
        while ((h + 1) * (h + 1) <= nucleus_radius) h++;
        hipLaunchKernelGGL(k_paint, dim3(grid_for(N)), dim3(NT), 0, s, P, d_flag, d_rank, n_grains, nucleus_radius,
                           h, d_grain, d_out, d_rex, d_nuc, d_vox, (long long)max_new);
        if (hipGetLastError() != hipSuccess) st = EVPGPU_ERR_HIP;
    }
    auto D2H = [&](void *h, const void *d, size_t b) {
        if (st == EVPGPU_OK && b && hipMemcpyAsync(h, d, b, hipMemcpyDeviceToHost, s) != hipSuccess)
            st = EVPGPU_ERR_HIP;
    };
    D2H(grain_out, d_out, N * sizeof(int));
    D2H(gid_rex, d_rex, N * sizeof(int));
    D2H(nuclei, d_nuc, N * sizeof(int));
    D2H(kc, d_kc, N * sizeof(double));
    D2H(rho_tot, d_rho, N * sizeof(double));
    D2H(nucleus_vox, d_vox, (size_t)(total < max_new ? total : max_new) * sizeof(long long));
    if (s) (void)hipStreamSynchronize(s);
    *n_new = (int)total;
    if (st == EVPGPU_OK && total > max_new) st = EVPGPU_ERR_ARG;
    void *all[] = {d_grain, d_out, d_flag, d_rex, d_nuc, d_r[0], d_r[1], d_r[2], d_r[3], d_r[4],
                   d_kappa, d_kc,   d_rho,  d_rank, d_vox, tmp};
    for (void *q : all) (void)hipFree(q);
    if (s) (void)hipStreamDestroy(s);
    return st;
}

extern "C" int evpgpu_drx_uniform_pos(unsigned long seed, long long n, double *out)
{
    if (n < 0 || !out) return EVPGPU_ERR_ARG;
    MT r(seed);
    for (long long v = 0; v < n; v++) out[v] = r.uniform_pos();
    return EVPGPU_OK;
}

extern "C" int evpgpu_drx_weibull(unsigned long seed, long long n, const double *a, double b, double *out)
{
    if (n < 0 || !a || !out || !(b != 0.0)) return EVPGPU_ERR_ARG;
    MT r(seed);
    for (long long v = 0; v < n; v++) {
        // gsl_ran_weibull: a * (-log x)^(1/b), x uniform on (0, 1)
        const double x = r.uniform_pos();
        const double z = pow(-log(x), 1 / b);
        out[v] = a[v] * z;
    }
    return EVPGPU_OK;
}
