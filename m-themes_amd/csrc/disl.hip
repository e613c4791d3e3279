// Dislocation-density law around the Green loop: the constitutive law of the
// build M-THEMES ships (src/evpFlags.h: DD_BASED_FLAG, DD_CU,
// FORWARD_PLUS_BACKWARD, DD_GND_FIRST_ORDER, PF_DRX before any DRX event).
//
// The rate law StrainRate_Orowan (src/constitutive.c:507-721) depends on the
// stress only through the resolved shears; its per-system constants -- the
// passing stress, 1/(tau_0 + tau_cut) and gam_o -- depend on the densities,
// which change once per increment (DislocationEvolution).  So:
//   k_dd_prologue  (once per increment, before the loop): forest and parallel
//       densities from the 12x12 interaction and projection tables, the
//       mobile density, and the constants the Newton kernel reads (RM_DISL);
//   k_dd_evolve    (once per increment, after the loop, before update_orient
//       turns the orientations): the shear-rate gradient with the grain-
//       boundary one-sided differences (Gradient_ShearRate,
//       src/constitutive.c:875-1108) fused with DislocationEvolution
//       (src/constitutive.c:1346-1520).
// Both are one thread per voxel, HBM-bound streams of the 12-per-voxel
// densities; the tables sit in LDS.
#include "evpgpu_impl.h"

namespace {

constexpr int NS = EVPGPU_NSYSMX;
constexpr int TAB = NS * NS * 9; // chi, then cos s,g1,g2,g3, sin s,g1,g2,g3
constexpr double TAU_0 = 1000;                 // src/constitutive.c:531
constexpr double TRIAL_FREQ_DRX = 1.0E10;      // src/evp.h:80
constexpr double MY_J_IN_EV = 0.62415063631;   // src/evp.h:83
constexpr double EPSILON_DD = 5.0E-3;          // src/PFgrid.h:50
constexpr double PI_REF = 3.14159265359;       // src/evp.h:77
constexpr int DD_THREADS = 256;

// Per crystal voxel: rho_F, rho_P (src/constitutive.c:560-573), then per
// system the passing stress with the mobile density of the previous call
// (the first call of the increment), the new mobile density, the passing
// stress every later call sees, 1/(tau_0 + tau_cut) and gam_o
// (src/constitutive.c:582-640).
__global__ __launch_bounds__(DD_THREADS) void k_dd_prologue(DDParams D, long long N,
                                                            const int *__restrict__ phase,
                                                            const double *__restrict__ tab,
                                                            double *__restrict__ rho,
                                                            double *__restrict__ kk)
{
    __shared__ double t[2 * TAB];
    for (int i = threadIdx.x; i < 2 * TAB; i += DD_THREADS) t[i] = tab[i];
    __syncthreads();
    const long long v = (long long)blockIdx.x * DD_THREADS + threadIdx.x;
    if (v >= N) return;
    const int p = phase[v] - 1;
    if (p < 0 || p > 1 || D.is_gas[p]) return;
    const double *chi = t + p * TAB;
    const double *cs = chi + 1 * NS * NS, *cg1 = chi + 2 * NS * NS, *cg2 = chi + 3 * NS * NS;
    const double *cg3 = chi + 4 * NS * NS, *ss = chi + 5 * NS * NS, *sg1 = chi + 6 * NS * NS;
    const double *sg2 = chi + 7 * NS * NS, *sg3 = chi + 8 * NS * NS;
    const long long S = NS * N;
    const double *rs = rho, *rg1 = rho + 2 * S, *rg2 = rho + 3 * S, *rg3 = rho + 4 * S;
    double *rm = rho + S, *rF = rho + 5 * S, *rP = rho + 6 * S;
    double a_s[NS], a_g1[NS], a_g2[NS], a_g3[NS];
#pragma unroll
    for (int j = 0; j < NS; j++) {
        a_s[j] = rs[j * N + v];
        a_g1[j] = rg1[j * N + v];
        a_g2[j] = rg2[j * N + v];
        a_g3[j] = rg3[j * N + v];
    }
    const int ns = D.nsys[p];
    const double c1 = D.c[p][0], c2 = D.c[p][1], c3 = D.c[p][2], G = D.shear_G[p];
    const double kT = D.kT, ex = exp(-1.0 * D.q_slip[p]);
    for (int i = 0; i < ns; i++) {
        double F = 0.0, P = 0.0;
        for (int j = 0; j < ns; j++) {
            const int ij = i * NS + j;
            F += chi[ij] * (a_s[j] * cs[ij] + fabs(a_g1[j] * cg1[ij]) + fabs(a_g2[j] * cg2[ij]) +
                            fabs(a_g3[j] * cg3[ij]));
            P += chi[ij] * (a_s[j] * ss[ij] + fabs(a_g1[j] * sg1[ij]) + fabs(a_g2[j] * sg2[ij]) +
                            fabs(a_g3[j] * sg3[ij]));
        }
        const double b = D.bbB[p][i];
        const double m_old = rm[i * N + v];
        const double tp_first = c1 * G * b * sqrt(fabs(P + m_old)) * 1.0E-3 + D.drx;
        const double tau_cut = (kT / MY_J_IN_EV) * sqrt(F) / c2 / c3 / b / b * 0.1;
        const double m = 2.0 * kT / c1 / c2 / c3 / G / (pow(b, 3.0)) * sqrt(P * F) * 160.2177;
        const double g0 = 1E-3 * m * TRIAL_FREQ_DRX * c2 * b * ex / sqrt(F);
        const double tp = c1 * G * b * sqrt(fabs(P + m)) * 1.0E-3 + D.drx;
        rF[i * N + v] = F;
        rP[i * N + v] = P;
        rm[i * N + v] = m;
        kk[(0 * NS + i) * N + v] = tp;
        kk[(1 * NS + i) * N + v] = 1.0 / (TAU_0 + tau_cut);
        kk[(2 * NS + i) * N + v] = g0;
        kk[(3 * NS + i) * N + v] = tp_first;
    }
}

// one-sided or central difference by the two grain checks
// (src/constitutive.c:1012-1053); 0 when both neighbours are other grains
__device__ __forceinline__ double gb_diff(bool same_p, bool same_m, double f, double fp, double fm)
{
    if (!same_p && same_m) return (f - fm);
    if (same_p && !same_m) return (fp - f);
    if (same_p && same_m) return (fp - fm) / 2.0;
    return 0.0;
}

// Gradient_ShearRate + DislocationEvolution of one crystal voxel.  The
// Schmid vectors are those of the loop (orientations before update_orient);
// the grain checks of Gradient_ExchangeGrainID (src/constitutive.c:818, 838,
// 858) compare the z-minus side with the -y neighbour.
__global__ __launch_bounds__(DD_THREADS) void k_dd_evolve(
    DDParams D, int nx, int ny, int nz, const int *__restrict__ phase, const int *__restrict__ grain,
    const double *__restrict__ sig, const double *__restrict__ tran, const double *__restrict__ gamdot,
    double *__restrict__ rho)
{
    const long long N = (long long)nx * ny * nz;
    const long long v = (long long)blockIdx.x * DD_THREADS + threadIdx.x;
    if (v >= N) return;
    const int p = phase[v] - 1;
    if (p < 0 || p > 1 || D.is_gas[p]) return;
    const int x = (int)(v / ((long long)ny * nz)), y = (int)((v / nz) % ny), z = (int)(v % nz);
    const long long S = NS * N;
    double *rs = rho, *rg1 = rho + 2 * S, *rg2 = rho + 3 * S, *rg3 = rho + 4 * S;
    double *rdip = rho + 7 * S, *rfor = rho + 8 * S;
    const double *rm = rho + S, *rF = rho + 5 * S;
    // stress in the B-basis (chg_basis opt 2), hydro as the reference forms it
    double s6[6];
    {
        double t[6];
#pragma unroll
        for (int c = 0; c < 6; c++) t[c] = sig[c * N + v];
        s6[0] = t[0] * (-1. * RSQ2) + t[1] * RSQ2;
        s6[1] = t[0] * (-1. * RSQ6) + t[1] * (-1. * RSQ6) + t[2] * (2. * RSQ6);
        s6[2] = t[3] * RSQ2 + t[3] * RSQ2;
        s6[3] = t[4] * RSQ2 + t[4] * RSQ2;
        s6[4] = t[5] * RSQ2 + t[5] * RSQ2;
        s6[5] = t[0] * RSQ3 + t[1] * RSQ3 + t[2] * RSQ3;
    }
    const double hydro = (s6[0] + s6[1] + s6[2]) / 3.0;
    double T[3][3];
#pragma unroll
    for (int q = 0; q < 9; q++) T[q / 3][q % 3] = tran[q * N + v];
    // neighbours and grain checks (periodic)
    long long vxp = ((long long)((x + 1) % nx) * ny + y) * nz + z;
    long long vxm = ((long long)((x + nx - 1) % nx) * ny + y) * nz + z;
    long long vyp = ((long long)x * ny + (y + 1) % ny) * nz + z;
    long long vym = ((long long)x * ny + (y + ny - 1) % ny) * nz + z;
    long long vzp = ((long long)x * ny + y) * nz + (z + 1) % nz;
    long long vzm = ((long long)x * ny + y) * nz + (z + nz - 1) % nz;
    const int g = grain[v];
    const bool kxp = g == grain[vxp], kxm = g == grain[vxm], kyp = g == grain[vyp],
               kym = g == grain[vym], kzp = g == grain[vzp], kzm = kym;
    const int ns = D.nsys[p];
    const double c4 = D.c[p][3], c5 = D.c[p][4], c6 = D.c[p][5], c7 = D.c[p][6];
    const double q_bulk = D.q_bulk[p] * MY_J_IN_EV / D.kT;
    const double G = D.shear_G[p], nu = D.nu[p], dt = D.dt;
    const double climb = exp(-1.0 * q_bulk);
    for (int i = 0; i < ns; i++) {
        // Schmid vector in sample axes, as k_newton forms it (update_schmid)
        double tn[3], tb[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            tn[a] = T[a][0] * D.bn[p][i][0] + T[a][1] * D.bn[p][i][1] + T[a][2] * D.bn[p][i][2];
            tb[a] = T[a][0] * D.bb[p][i][0] + T[a][1] * D.bb[p][i][1] + T[a][2] * D.bb[p][i][2];
        }
        const double P00 = tn[0] * tb[0], P11 = tn[1] * tb[1], P22 = tn[2] * tb[2];
        const double P12 = (tn[1] * tb[2] + tn[2] * tb[1]) / 2.0;
        const double P02 = (tn[0] * tb[2] + tn[2] * tb[0]) / 2.0;
        const double P01 = (tn[0] * tb[1] + tn[1] * tb[0]) / 2.0;
        const double sc0 = P00 * (-1. * RSQ2) + P11 * RSQ2;
        const double sc1 = P00 * (-1. * RSQ6) + P11 * (-1. * RSQ6) + P22 * (2. * RSQ6);
        const double sc2 = P12 * RSQ2 + P12 * RSQ2;
        const double sc3 = P02 * RSQ2 + P02 * RSQ2;
        const double sc4 = P01 * RSQ2 + P01 * RSQ2;
        const double rss = sc0 * s6[0] + sc1 * s6[1] + sc2 * s6[2] + sc3 * s6[3] + sc4 * s6[4];
        const double b = D.bbB[p][i];
        const double d_dipole = sqrt(3.0) * G * b / 16.0 / PI_REF / (1 - nu) / (fabs(rss) + EPSILON_DD);
        const long long iv = (long long)i * N + v;
        const double gd = gamdot[iv];
        if (D.istep <= 2000000) {
            const double r = rs[iv];
            // dipole and forest densities formed by slip (diagnostic sums)
            rdip[iv] = rdip[iv] + fabs(gd) * (c6 * (d_dipole)*rm[iv]) * dt;
            rfor[iv] = rfor[iv] + fabs(gd) * (c4 * sqrt(rF[iv])) * dt;
            const double dot_ssd =
                fabs(gd) * (c4 * sqrt(rF[iv]) + c6 * (d_dipole)*rm[iv] - c5 * r) -
                c7 * 1.6 * 1E12 * 1E-9 * r * r -
                4 * 1E12 * (d_dipole) * (D.D_0 / b) * pow(r, 2.0) *
                    exp((fabs(hydro) * b * b * b * 100 / (1.38 * D.T_K)) - 1.0) * climb;
            rs[iv] = r + dot_ssd * dt;
        }
        if (D.istep <= 200000000) {
            double gx = 0.0, gy = 0.0, gz = 0.0;
            if (D.use_grad) {
                const double *f = gamdot + (long long)i * N;
                gx = gb_diff(kxp, kxm, gd, f[vxp], f[vxm]);
                gy = gb_diff(kyp, kym, gd, f[vyp], f[vym]);
                gz = gb_diff(kzp, kzm, gd, f[vzp], f[vzm]);
            }
            const double *bp = D.bp[p][i], *bb = D.bb[p][i];
            double gnd1 = (gx * bp[0] + gy * bp[1] + gz * bp[2]);
            rg1[iv] = rg1[iv] + gnd1 / b / D.L0 * dt;
            gnd1 = (gx * bb[0] + gy * bb[1] + gz * bb[2]);
            rg2[iv] = rg2[iv] + -1.0 * gnd1 / b / D.L0 * dt;
            rg3[iv] = rg3[iv] + 0.0 * dt;
        }
    }
}

// AvgDislocation (src/constitutive.c:1669-1780): per block, the sums over
// its crystal voxels and their systems of rho_F, rho_P, rho_s, rho_m,
// |(rho_g1, rho_g2, rho_g3)|, rho_dipole and rho_forest (fixed-order tree)
constexpr int NAVG = 7;
struct DdAvgPhases {
    int is_gas[2], nsys[2];
};
__global__ __launch_bounds__(DD_THREADS) void k_dd_avg(long long N, const int *__restrict__ phase,
                                                       DdAvgPhases ph, const double *__restrict__ rho,
                                                       double *__restrict__ partials)
{
    __shared__ double sh[NAVG][DD_THREADS];
    const long long v = (long long)blockIdx.x * DD_THREADS + threadIdx.x;
    double a[NAVG] = {0, 0, 0, 0, 0, 0, 0};
    if (v < N) {
        const int p = phase[v] - 1;
        if (p >= 0 && p <= 1 && !ph.is_gas[p]) {
            const long long S = NS * N;
            for (int i = 0; i < ph.nsys[p]; i++) {
                const long long iv = (long long)i * N + v;
                const double g1 = rho[2 * S + iv], g2 = rho[3 * S + iv], g3 = rho[4 * S + iv];
                a[0] += rho[5 * S + iv];
                a[1] += rho[6 * S + iv];
                a[2] += rho[iv];
                a[3] += rho[S + iv];
                a[4] += sqrt(g1 * g1 + g2 * g2 + g3 * g3);
                a[5] += rho[7 * S + iv];
                a[6] += rho[8 * S + iv];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NAVG; k++) sh[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int w = DD_THREADS / 2; w > 0; w /= 2) {
        if (threadIdx.x < w)
#pragma unroll
            for (int k = 0; k < NAVG; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < NAVG) partials[(long long)blockIdx.x * NAVG + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_fill_value(double *p, long long n, double v)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

} // namespace

// kT of src/evp.h:86
static double dd_kT(double T_K) { return (T_K * 1.380658e-23 * (1.0 / 1.60217733e-19)); }

DDParams dd_params(const evpgpu_ctx *c)
{
    DDParams D;
    memset(&D, 0, sizeof(D));
    D.kT = dd_kT(c->dd_law.T_K);
    D.T_K = c->dd_law.T_K;
    D.L0 = c->dd_law.L0;
    D.D_0 = c->dd_law.D_0;
    D.dt = c->ld.dt;
    // the DRX term of the passing stress (src/constitutive.c:584-586) with
    // t_last at its initial -1E3 tau_DRX TimeStep (src/init.c:1025)
    const double t_last = -1E3 * c->dd_law.tau_drx * c->ld.dt;
    D.drx = c->dd_law.delta_tau_drx * exp((t_last - c->m.time_tot) / c->ld.dt / c->dd_law.tau_drx);
    for (int p = 0; p < 2; p++) {
        D.is_gas[p] = p < c->nphases ? c->ph[p].is_gas : 1;
        if (D.is_gas[p]) continue;
        const evpgpu_dd_phase &q = c->dd_ph[p];
        memcpy(D.c[p], q.c, sizeof(D.c[p]));
        D.shear_G[p] = q.shear_G;
        D.nu[p] = q.poisson_nu;
        D.q_slip[p] = q.q_slip * MY_J_IN_EV / D.kT;
        D.q_bulk[p] = q.q_bulk;
        memcpy(D.bbB[p], q.bbB, sizeof(D.bbB[p]));
        memcpy(D.bp[p], q.bp, sizeof(D.bp[p]));
        memcpy(D.bn[p], c->ph[p].bn, sizeof(D.bn[p]));
        memcpy(D.bb[p], c->ph[p].bb, sizeof(D.bb[p]));
        D.nsys[p] = c->ph[p].nsys;
    }
    D.istep = c->m.istep;
    D.use_grad = c->ld.hard_flag == 1 && c->m.istep > 1 && c->m.istep <= 20000000;
    return D;
}

int dd_prologue_launch(evpgpu_ctx *c, const DDParams &D)
{
    const long long nb = (c->N + DD_THREADS - 1) / DD_THREADS;
    hipLaunchKernelGGL(k_dd_prologue, dim3((unsigned)nb), dim3(DD_THREADS), 0, c->stream, D, c->N,
                       c->phase, c->dd_tab, c->rho, c->dd_k);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

int dd_evolve_launch(evpgpu_ctx *c, const DDParams &D)
{
    const long long nb = (c->N + DD_THREADS - 1) / DD_THREADS;
    hipLaunchKernelGGL(k_dd_evolve, dim3((unsigned)nb), dim3(DD_THREADS), 0, c->stream, D, c->nx, c->ny,
                       c->nz, c->phase, c->grain, c->sig, c->tran, c->gamdot, c->rho);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

extern "C" {

int evpgpu_set_dd_law(evpgpu_ctx *c, const evpgpu_dd_law *law, const evpgpu_dd_phase *phases, int nphases)
{
    if (!c || !law || !phases || nphases != c->nphases) return EVPGPU_ERR_ARG;
    if (!c->have_phases) return EVPGPU_ERR_STATE;
    for (int p = 0; p < nphases; p++)
        if (!c->ph[p].is_gas && c->ph[p].nsys != NS) return EVPGPU_ERR_ARG;
    const long long N = c->N;
    if (!c->rho && hipMalloc((void **)&c->rho, sizeof(double) * 9 * NS * N) != hipSuccess) {
        c->rho = nullptr;
        return EVPGPU_ERR_NOMEM;
    }
    if (!c->dd_k && hipMalloc((void **)&c->dd_k, sizeof(double) * 4 * NS * N) != hipSuccess) {
        c->dd_k = nullptr;
        return EVPGPU_ERR_NOMEM;
    }
    if (!c->dd_tab && hipMalloc((void **)&c->dd_tab, sizeof(double) * 2 * TAB) != hipSuccess) {
        c->dd_tab = nullptr;
        return EVPGPU_ERR_NOMEM;
    }
    c->dd_law = *law;
    memset(c->dd_ph, 0, sizeof(c->dd_ph));
    double *tab = (double *)calloc(2 * TAB, sizeof(double));
    if (!tab) return EVPGPU_ERR_NOMEM;
    for (int p = 0; p < nphases; p++) {
        c->dd_ph[p] = phases[p];
        memcpy(tab + p * TAB, phases[p].chi, sizeof(phases[p].chi));
        memcpy(tab + p * TAB + NS * NS, phases[p].cth, sizeof(phases[p].cth));
    }
    int st = EVPGPU_OK;
    if (hipMemcpy(c->dd_tab, tab, sizeof(double) * 2 * TAB, hipMemcpyHostToDevice) != hipSuccess)
        st = EVPGPU_ERR_HIP;
    free(tab);
    // initial densities (src/init.c:1073-1093): rho_s of every voxel and
    // system, the others 0
    if (!st) {
        hipLaunchKernelGGL(k_fill_value, dim3((unsigned)((NS * N + 255) / 256)), dim3(256), 0, c->stream, c->rho,
                           NS * N, law->rho_ssd_initial);
        if (hipGetLastError() != hipSuccess) st = EVPGPU_ERR_HIP;
    }
    if (!st) st = fill_zero(c, c->rho + NS * N, 8 * NS * N);
    if (!st) st = fill_zero(c, c->dd_k, 4 * NS * N);
    if (!st && hipStreamSynchronize(c->stream) != hipSuccess) st = EVPGPU_ERR_HIP;
    if (!st) c->dd_on = 1;
    return st;
}

int evpgpu_dd_averages(evpgpu_ctx *c, double avg[7])
{
    if (!c || !avg) return EVPGPU_ERR_ARG;
    if (!c->dd_on) return EVPGPU_ERR_STATE;
    DdAvgPhases h;
    for (int p = 0; p < 2; p++) {
        h.is_gas[p] = p < c->nphases ? c->ph[p].is_gas : 1;
        h.nsys[p] = p < c->nphases ? c->ph[p].nsys : 0;
    }
    int st = EVPGPU_OK;
    const int nb = (int)((c->N + DD_THREADS - 1) / DD_THREADS);
    hipLaunchKernelGGL(k_dd_avg, dim3(nb), dim3(DD_THREADS), 0, c->stream, c->N, c->phase, h, c->rho,
                       c->partials);
    if (hipGetLastError() != hipSuccess) st = EVPGPU_ERR_HIP;
    if (!st) st = reduce_partials(c, c->partials, nb, NAVG, c->red);
    double s[NAVG];
    if (!st && hipMemcpyAsync(s, c->red, sizeof(s), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        st = EVPGPU_ERR_HIP;
    if (!st && hipStreamSynchronize(c->stream) != hipSuccess) st = EVPGPU_ERR_HIP;
    if (st) return st;
    for (int k = 0; k < NAVG; k++) avg[k] = s[k] * (1.0 / (double)c->N);
    return EVPGPU_OK;
}

int evpgpu_download_dd_state(evpgpu_ctx *c, double *rho_s, double *rho_m, double *rho_g1, double *rho_g2,
                             double *rho_g3, double *rho_f, double *rho_p)
{
    if (!c) return EVPGPU_ERR_ARG;
    if (!c->dd_on) return EVPGPU_ERR_STATE;
    const long long N = c->N;
    double *outs[7] = {rho_s, rho_m, rho_g1, rho_g2, rho_g3, rho_f, rho_p};
    double *tmp = (double *)malloc(sizeof(double) * NS * N);
    if (!tmp) return EVPGPU_ERR_NOMEM;
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        free(tmp);
        return EVPGPU_ERR_HIP;
    }
    for (int k = 0; k < 7; k++) {
        if (!outs[k]) continue;
        if (hipMemcpy(tmp, c->rho + (long long)k * NS * N, sizeof(double) * NS * N, hipMemcpyDeviceToHost) !=
            hipSuccess) {
            free(tmp);
            return EVPGPU_ERR_HIP;
        }
        // [system][voxel] -> [voxel][system]
        for (long long v = 0; v < N; v++)
            for (int s = 0; s < NS; s++) outs[k][v * NS + s] = tmp[(long long)s * N + v];
    }
    free(tmp);
    return EVPGPU_OK;
}

} // extern "C"
