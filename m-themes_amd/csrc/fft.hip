// fft.hip -- Stockham 3-D R2C / C2R FFT for gfx950, FP64 and FP32.
//
// Replaces FFTW-MPI's fftwnd_mpi (src/evolution.c:1367,1399; plans at
// src/init.c:752-757) with whole-grid-resident transforms.  The reference runs
// 9 full complex transforms each way per Green iteration on real data; here
// the stress components go through R2C along z (half spectrum kz = 0..nz/2),
// then C2C along y and x, and come back through C2R, which gives
// Re(IFFT(.)) of the Hermitian spectrum the Green step produces.
//
// Each pass moves every line of one axis through LDS:
//   * a workgroup owns TB = 16 lines whose elements are adjacent in memory
//     (consecutive kz for the strided y/x passes, consecutive rows for z), so
//     every global load/store is a run of TB complex values;
//   * LDS rows are padded to L+1 entries so the transposed accesses of the
//     load/store phases spread over the banks;
//   * the butterflies are Stockham autosort radix-8 stages (radix-4/2 last),
//     each thread holding whole butterflies in registers, one LDS round trip
//     per stage, no bit reversal.
// Twiddles come from per-length tables computed on the host in double.
#include <math.h>

#include "evpgpu_impl.h"

#include "fft_core.h"

namespace {

using namespace fftc;

constexpr int TB = 16;      // lines per workgroup
constexpr int NT = 256;     // threads per workgroup

// One Stockham stage over TB lines of length L held in LDS (pitch L+1).
template <bool INV, int L, int RAD, int NS, typename C>
__device__ __forceinline__ void stage(C *lds, const C *tw)
{
    constexpr int P = L + 1;
    constexpr int NB = L / RAD;                 // butterflies per line
    constexpr int TOT = NB * TB;
    constexpr int PER = (TOT + NT - 1) / NT;    // butterflies per thread
    C v[PER][RAD];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        int b = threadIdx.x + q * NT;
        if (b < TOT) {
            int t = b / NB, j = b % NB;
#pragma unroll
            for (int r = 0; r < RAD; r++) v[q][r] = lds[t * P + j + r * NB];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++) {
        int b = threadIdx.x + q * NT;
        if (b < TOT) {
            int t = b / NB, j = b % NB;
            int k = j % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < RAD; r++) {
                    C w = tw[(r * k * (L / (NS * RAD))) % L];
                    if (INV) w.y = -w.y;
                    v[q][r] = cmul(v[q][r], w);
                }
            }
            bfly<INV, RAD>(v[q]);
            int d = (j / NS) * NS * RAD + k;
#pragma unroll
            for (int r = 0; r < RAD; r++) lds[t * P + d + r * NS] = v[q][r];
        }
    }
    __syncthreads();
}

template <bool INV, int L, int S, int NS, typename C>
__device__ __forceinline__ void stages(C *lds, const C *tw)
{
    if constexpr (S < nstages(L)) {
        constexpr int RAD = radix_at(L, S);
        stage<INV, L, RAD, NS>(lds, tw);
        stages<INV, L, S + 1, NS * RAD>(lds, tw);
    }
}

// Workgroups go to the 8 XCDs round robin (block b to XCD b mod 8), each
// XCD with its own L2.  Tiles t and t + 1 share the 128-byte lines at their
// boundary, so within each group of 16 blocks block g*16 + r takes tile
// g*16 + 2 (r mod 8) + r / 8: the pair lands on blocks b and b + 8, one XCD.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    if (b >= nb / 16 * 16) return b;
    const int g = b / 16, r = b % 16;
    return g * 16 + (r % 8) * 2 + r / 8;
}

// ---- strided C2C pass: lines along an axis with element stride `es`.
// Tile (bx, o) of nbx x nout (1-D grid, XCD-aware order): lines inner =
// bx*TB .. +TB of group o, group base (o / O2) * S1 + (o % O2) * S2.
template <typename R, int L, bool INV>
__global__ __launch_bounds__(NT) void k_fft_strided(typename CT<R>::T *data, int ninner, int nbx,
                                                    long long es, int O2, long long S1,
                                                    long long S2, const typename CT<R>::T *tw)
{
    typedef typename CT<R>::T C;
    constexpr int P = L + 1;
    __shared__ C lds[TB * P];
    // the twiddle table in LDS, loaded with the tile (from memory, every
    // stage waited on its twiddles)
    __shared__ C s_tw[L];
    for (int i = threadIdx.x; i < L; i += NT) s_tw[i] = tw[i];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int t0 = (tile % nbx) * TB;
    const int o = tile / nbx;
    const long long base = (long long)(o / O2) * S1 + (long long)(o % O2) * S2 + t0;
    const int nt = min(TB, ninner - t0);
    staged_copy<NT, (L * TB + NT - 1) / NT, sizeof(R) == 8 ? 8 : 16, C>(
        L * TB,
        [&](int e) {
            const int i = e / TB, t = e % TB;
            return t < nt ? data[base + i * es + t] : C{};
        },
        [&](int e, C v) {
            const int i = e / TB, t = e % TB;
            if (t < nt) lds[t * P + i] = v;
        });
    __syncthreads();
    stages<INV, L, 0, 1>(lds, (const C *)s_tw);
    for (int e = threadIdx.x; e < L * TB; e += NT) {
        int i = e / TB, t = e % TB;
        if (t < nt) data[base + i * es + t] = lds[t * P + i];
    }
}

// ---- z pass, real -> half spectrum (ncomp x N real -> ncomp x NK complex):
// TB lines per workgroup, each transformed as an L/2-point complex FFT of
// the packed pairs and split into its L/2 + 1 half-spectrum values.
// IN 0: FP64 field; IN 1: FP64 field minus an FP32 anchor field (the mixed
// loop's delta); IN 2: an FP32 field.
template <typename R, int L, int IN>
__global__ __launch_bounds__(NT) void k_fft_z_r2c(const double *in, const float *inA,
                                                  typename CT<R>::T *out, long long N,
                                                  long long NK, const typename CT<R>::T *tw)
{
    typedef typename CT<R>::T C;
    constexpr int M = L / 2, P = M + 1;
    __shared__ C lds[TB * P];
    __shared__ C s_tw[L]; // the twiddle table, loaded with the lines
    for (int i = threadIdx.x; i < L; i += NT) s_tw[i] = tw[i];
    const int comp = blockIdx.y;
    const long long line0 = (long long)blockIdx.x * TB;
    const long long off = comp * N + line0 * L; // even: the pairs below are aligned
    staged_copy<NT, (M * TB + NT - 1) / NT, IN == 1 ? 8 : 16, C>(
        M * TB,
        [&](int e) {
            R a, b;
            if (IN == 0) {
                double2 v = reinterpret_cast<const double2 *>(in + off)[e];
                a = (R)v.x;
                b = (R)v.y;
            } else if (IN == 1) {
                double2 v = reinterpret_cast<const double2 *>(in + off)[e];
                float2 w = reinterpret_cast<const float2 *>(inA + off)[e];
                a = (R)(v.x - (double)w.x);
                b = (R)(v.y - (double)w.y);
            } else {
                float2 w = reinterpret_cast<const float2 *>(inA + off)[e];
                a = (R)w.x;
                b = (R)w.y;
            }
            return C{a, b};
        },
        [&](int e, C v) { lds[(e / M) * P + e % M] = v; });
    __syncthreads();
    // the M-point transforms: the length-L table read at stride 2
    gstages<false, M, 0, 1, TB, P, 1, 0, NT, 2>(lds, (const C *)s_tw);
    for (int i = threadIdx.x; i < TB * (M / 2 + 1); i += NT)
        split_forward<R, M>(lds + (i / (M / 2 + 1)) * P, i % (M / 2 + 1), (const C *)s_tw);
    __syncthreads();
    C *dst = out + comp * NK + line0 * P;
    for (int e = threadIdx.x; e < P * TB; e += NT) dst[e] = lds[e];
}

// ---- z pass, half spectrum -> real, Re(IFFT) of the Hermitian extension
// (the imaginary parts of k = 0 and k = L/2 dropped), through the inverse
// split and an L/2-point complex FFT.
// MODE 0: target[c] += add[c] + re*scale     (the strain update of
//         src/evolution.c:1400-1401)
// MODE 1: target[c]  = re*scale
// MODE 2: target[c] += add[c] + fa[c] + re*scale   (mixed loop: anchor + delta)
// MODE 3: fo[c]      = (float)(re*scale)           (mixed loop: anchor update)
// MODE 4: target[c]  = base[c] + re*scale          (projection of a strain change)
template <typename R, int L, int MODE>
__global__ __launch_bounds__(NT) void k_fft_z_c2r(const typename CT<R>::T *in, double *target,
                                                  const float *fa, float *fo, const double *base,
                                                  long long N, long long NK, double a0, double a1,
                                                  double a2, double a3, double a4, double a5,
                                                  double scale, const typename CT<R>::T *tw)
{
    typedef typename CT<R>::T C;
    constexpr int M = L / 2, P = M + 1;
    __shared__ C lds[TB * P];
    __shared__ C s_tw[L]; // the twiddle table, loaded with the lines
    for (int i = threadIdx.x; i < L; i += NT) s_tw[i] = tw[i];
    const int comp = blockIdx.y;
    const long long line0 = (long long)blockIdx.x * TB;
    const C *src = in + comp * NK + line0 * P;
    staged_copy<NT, (P * TB + NT - 1) / NT, sizeof(R) == 8 ? 8 : 16, C>(
        P * TB, [&](int e) { return src[e]; }, [&](int e, C v) { lds[e] = v; });
    __syncthreads();
    for (int i = threadIdx.x; i < TB * (M / 2 + 1); i += NT)
        split_inverse<R, M>(lds + (i / (M / 2 + 1)) * P, i % (M / 2 + 1), (const C *)s_tw);
    __syncthreads();
    gstages<true, M, 0, 1, TB, P, 1, 0, NT, 2>(lds, (const C *)s_tw);
    const double add = comp == 0 ? a0 : comp == 1 ? a1 : comp == 2 ? a2 : comp == 3 ? a3 : comp == 4 ? a4 : a5;
    const long long off = comp * N + line0 * L;
    for (int e = threadIdx.x; e < M * TB; e += NT) {
        const int t = e / M, n = e % M;
        const C z = lds[t * P + n];
        const double r0 = (double)z.x * scale, r1 = (double)z.y * scale;
        double2 *t2 = reinterpret_cast<double2 *>(target + off);
        if (MODE == 0) {
            double2 v = t2[e];
            t2[e] = make_double2(v.x + (add + r0), v.y + (add + r1));
        } else if (MODE == 1) {
            t2[e] = make_double2(r0, r1);
        } else if (MODE == 2) {
            double2 v = t2[e];
            float2 f = reinterpret_cast<const float2 *>(fa + off)[e];
            t2[e] = make_double2(v.x + (add + (double)f.x + r0), v.y + (add + (double)f.y + r1));
        } else if (MODE == 3) {
            reinterpret_cast<float2 *>(fo + off)[e] = make_float2((float)r0, (float)r1);
        } else {
            double2 b = reinterpret_cast<const double2 *>(base + off)[e];
            t2[e] = make_double2(b.x + r0, b.y + r1);
        }
    }
}

// ---- any length: two-level DFTs.  An axis whose length is not a power of
// two (FFTW, which the reference calls, takes any length) is transformed with
// L = L1 L2 (L1 the largest divisor <= sqrt L): length-L1 DFTs down the
// columns of the L1 x L2 view of each line, the twiddles w^(n2 k1), then
// length-L2 DFTs along the rows (Cooley-Tukey with direct inner DFTs), the
// lines of a workgroup and the twiddle table exp(-2 pi i m / L) in LDS.
// O(L1 + L2) work per element; a prime length is one direct DFT.  Same
// output layout and modes as the Stockham kernels above.
constexpr int DFT_LDS = 4096; // complex values per workgroup: 64 KiB in FP64

__host__ __device__ constexpr int dft_lines(int L)
{
    return (DFT_LDS / L - 1) / 2 < TB ? (DFT_LDS / L - 1) / 2 : TB;
}

__host__ __device__ inline int dft_l1(int L)
{
    int l1 = 1;
    for (int d = 2; d * d <= L; d++)
        if (L % d == 0) l1 = d;
    return l1;
}

// the TBL lines of a (TBL x L, complex) -> out(t, k) for k < KOUT, via b
// (TBL x L scratch); tw: the length-L table in LDS.  OUT(t, k, value)
// stores one output.
template <bool INV, typename C, typename OUT>
__device__ __forceinline__ void dft2_lines(const C *a, C *b, const C *tw, int L, int TBL, int nt, int KOUT,
                                           OUT out)
{
    const int L1 = dft_l1(L), L2 = L / L1;
    // stage A: b[t][k1*L2 + n2] = w^(n2 k1) sum_n1 a[t][n1*L2 + n2] w^(n1 k1 L2)
    for (int e = threadIdx.x; e < L * TBL; e += NT) {
        const int t = e / L, r = e % L;
        if (t >= nt) continue;
        const int k1 = r / L2, n2 = r % L2;
        const C *x = a + t * L + n2;
        const int step = (k1 * L2) % L;
        decltype(C{}.x) re = 0, im = 0;
        int m = 0;
        for (int n1 = 0; n1 < L1; n1++) {
            const C w = tw[m], v = x[n1 * L2];
            const auto wy = INV ? -w.y : w.y;
            re += v.x * w.x - v.y * wy;
            im += v.x * wy + v.y * w.x;
            m += step;
            if (m >= L) m -= L;
        }
        const C w = tw[(n2 * k1) % L];
        const auto wy = INV ? -w.y : w.y;
        b[t * L + r] = C{re * w.x - im * wy, re * wy + im * w.x};
    }
    __syncthreads();
    // stage B: X[k1 + L1 k2] = sum_n2 b[t][k1*L2 + n2] w^(n2 k2 L1)
    for (int e = threadIdx.x; e < KOUT * TBL; e += NT) {
        const int t = e / KOUT, k = e % KOUT;
        if (t >= nt) continue;
        const int k1 = k % L1, k2 = k / L1;
        const C *y = b + t * L + k1 * L2;
        const int step = (k2 * L1) % L;
        decltype(C{}.x) re = 0, im = 0;
        int m = 0;
        for (int n2 = 0; n2 < L2; n2++) {
            const C w = tw[m], v = y[n2];
            const auto wy = INV ? -w.y : w.y;
            re += v.x * w.x - v.y * wy;
            im += v.x * wy + v.y * w.x;
            m += step;
            if (m >= L) m -= L;
        }
        out(t, k, C{re, im});
    }
}

// strided C2C along an axis of length L: the lines of one group, as
// k_fft_strided tiles them (TBL = dft_lines(L) lines per workgroup)
template <typename R, bool INV>
__global__ __launch_bounds__(NT) void k_dft_strided(typename CT<R>::T *data, int L, int ninner, int nbx,
                                                    long long es, int O2, long long S1, long long S2,
                                                    const typename CT<R>::T *tw)
{
    typedef typename CT<R>::T C;
    extern __shared__ unsigned char dft_smem[];
    const int TBL = dft_lines(L);
    C *a = reinterpret_cast<C *>(dft_smem), *b = a + TBL * L, *tws = b + TBL * L;
    const int t0 = (blockIdx.x % nbx) * TBL;
    const int o = blockIdx.x / nbx;
    const long long base = (long long)(o / O2) * S1 + (long long)(o % O2) * S2 + t0;
    const int nt = min(TBL, ninner - t0);
    for (int i = threadIdx.x; i < L; i += NT) tws[i] = tw[i];
    for (int e = threadIdx.x; e < L * TBL; e += NT) {
        const int i = e / TBL, t = e % TBL;
        if (t < nt) a[t * L + i] = data[base + i * es + t];
    }
    __syncthreads();
    dft2_lines<INV>(a, b, tws, L, TBL, nt, L,
                    [&](int t, int k, C v) { data[base + k * es + t] = v; });
}

// real lines along z -> half spectrum (IN as k_fft_z_r2c)
template <typename R, int IN>
__global__ __launch_bounds__(NT) void k_dft_z_r2c(const double *in, const float *inA, typename CT<R>::T *out,
                                                  int L, long long nlines, long long N, long long NK,
                                                  const typename CT<R>::T *tw)
{
    typedef typename CT<R>::T C;
    extern __shared__ unsigned char dft_smem[];
    const int TBL = dft_lines(L), P = L / 2 + 1;
    C *a = reinterpret_cast<C *>(dft_smem), *b = a + TBL * L, *tws = b + TBL * L;
    const int comp = blockIdx.y;
    const long long line0 = (long long)blockIdx.x * TBL;
    const int nt = (int)min((long long)TBL, nlines - line0);
    const long long off = comp * N + line0 * L;
    for (int i = threadIdx.x; i < L; i += NT) tws[i] = tw[i];
    for (int e = threadIdx.x; e < L * TBL; e += NT) {
        if (e / L >= nt) continue;
        R v;
        if (IN == 0) v = (R)in[off + e];
        else if (IN == 1) v = (R)(in[off + e] - (double)inA[off + e]);
        else v = (R)inA[off + e];
        a[e] = C{v, (R)0};
    }
    __syncthreads();
    C *dst = out + comp * NK + line0 * P;
    dft2_lines<false>(a, b, tws, L, TBL, nt, P, [&](int t, int k, C v) { dst[t * P + k] = v; });
}

// half spectrum -> real lines along z: Re(IFFT) of the Hermitian extension
// (imaginary parts of k = 0 and, for even L, k = L/2 dropped), then MODE as
// k_fft_z_c2r
template <typename R, int MODE>
__global__ __launch_bounds__(NT) void k_dft_z_c2r(const typename CT<R>::T *in, double *target, const float *fa,
                                                  float *fo, const double *base, int L, long long nlines,
                                                  long long N, long long NK, double a0, double a1, double a2,
                                                  double a3, double a4, double a5, double scale,
                                                  const typename CT<R>::T *tw)
{
    typedef typename CT<R>::T C;
    extern __shared__ unsigned char dft_smem[];
    const int TBL = dft_lines(L), P = L / 2 + 1;
    C *a = reinterpret_cast<C *>(dft_smem), *b = a + TBL * L, *tws = b + TBL * L;
    const int comp = blockIdx.y;
    const long long line0 = (long long)blockIdx.x * TBL;
    const int nt = (int)min((long long)TBL, nlines - line0);
    const C *src = in + comp * NK + line0 * P;
    for (int i = threadIdx.x; i < L; i += NT) tws[i] = tw[i];
    for (int e = threadIdx.x; e < L * TBL; e += NT) {
        const int t = e / L, k = e % L;
        if (t >= nt) continue;
        C v;
        if (k < P) {
            v = src[t * P + k];
            if (k == 0 || 2 * k == L) v.y = (R)0;
        } else {
            v = src[t * P + (L - k)];
            v.y = -v.y;
        }
        a[e] = v;
    }
    __syncthreads();
    const double add = comp == 0 ? a0 : comp == 1 ? a1 : comp == 2 ? a2 : comp == 3 ? a3 : comp == 4 ? a4 : a5;
    const long long off = comp * N + line0 * L;
    dft2_lines<true>(a, b, tws, L, TBL, nt, L, [&](int t, int n, C z) {
        const long long i = off + (long long)t * L + n;
        const double r = (double)z.x * scale;
        if (MODE == 0) target[i] += add + r;
        else if (MODE == 1) target[i] = r;
        else if (MODE == 2) target[i] += add + (double)fa[i] + r;
        else if (MODE == 3) fo[i] = (float)r;
        else target[i] = base[i] + r;
    });
}

// lengths with a Stockham plan (2^m 3^a 5^b, divisible by 4, in [8, 256];
// the list of FFT_DISPATCH); the others take the two-level DFT kernels
bool is_stockham_len(int L)
{
    static const int lens[] = {8, 12, 16, 20, 24, 32, 36, 40, 48, 60, 64, 72, 80, 96, 100, 120,
                               128, 144, 160, 192, 200, 240, 256};
    for (int l : lens)
        if (l == L) return true;
    return false;
}

template <typename R> const typename CT<R>::T *twp(evpgpu_ctx *c, int axis);
template <> const double2 *twp<double>(evpgpu_ctx *c, int axis) { return c->tw64[axis]; }
template <> const float2 *twp<float>(evpgpu_ctx *c, int axis) { return c->tw32[axis]; }

// dispatch on a runtime length
#define FFT_DISPATCH(L_, CALL) \
    switch (L_) {              \
    case 8: CALL(8); break;    \
    case 12: CALL(12); break;  \
    case 16: CALL(16); break;  \
    case 20: CALL(20); break;  \
    case 24: CALL(24); break;  \
    case 32: CALL(32); break;  \
    case 36: CALL(36); break;  \
    case 40: CALL(40); break;  \
    case 48: CALL(48); break;  \
    case 60: CALL(60); break;  \
    case 64: CALL(64); break;  \
    case 72: CALL(72); break;  \
    case 80: CALL(80); break;  \
    case 96: CALL(96); break;  \
    case 100: CALL(100); break; \
    case 120: CALL(120); break; \
    case 128: CALL(128); break; \
    case 144: CALL(144); break; \
    case 160: CALL(160); break; \
    case 192: CALL(192); break; \
    case 200: CALL(200); break; \
    case 240: CALL(240); break; \
    case 256: CALL(256); break; \
    default: return EVPGPU_ERR_SIZE; \
    }

template <typename R, bool INV>
int pass_y(evpgpu_ctx *c, typename CT<R>::T *d, int ncomp)
{
    if (!is_stockham_len(c->ny)) {
        const int L = c->ny, tbl = dft_lines(L), nbx = (c->nzh + tbl - 1) / tbl;
        hipLaunchKernelGGL((k_dft_strided<R, INV>), dim3(nbx * ncomp * c->nx), dim3(NT),
                           sizeof(typename CT<R>::T) * (2 * tbl + 1) * L, c->stream, d, L, c->nzh, nbx, (long long)c->nzh, 1,
                           (long long)c->ny * c->nzh, 0LL, twp<R>(c, 1));
        return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
    }
    const int nbx = (c->nzh + TB - 1) / TB;
    dim3 g(nbx * ncomp * c->nx);
#define CALLY(L) hipLaunchKernelGGL((k_fft_strided<R, L, INV>), g, dim3(NT), 0, c->stream, d, c->nzh, nbx, \
                                    (long long)c->nzh, 1, (long long)c->ny * c->nzh, 0LL, twp<R>(c, 1))
    FFT_DISPATCH(c->ny, CALLY);
#undef CALLY
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

template <typename R, bool INV>
int pass_x(evpgpu_ctx *c, typename CT<R>::T *d, int ncomp)
{
    if (!is_stockham_len(c->nx)) {
        const int L = c->nx, tbl = dft_lines(L), nbx = (c->nzh + tbl - 1) / tbl;
        hipLaunchKernelGGL((k_dft_strided<R, INV>), dim3(nbx * ncomp * c->ny), dim3(NT),
                           sizeof(typename CT<R>::T) * (2 * tbl + 1) * L, c->stream, d, L, c->nzh, nbx,
                           (long long)c->ny * c->nzh, c->ny, (long long)c->NK, (long long)c->nzh, twp<R>(c, 0));
        return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
    }
    const int nbx = (c->nzh + TB - 1) / TB;
    dim3 g(nbx * ncomp * c->ny);
#define CALLX(L) hipLaunchKernelGGL((k_fft_strided<R, L, INV>), g, dim3(NT), 0, c->stream, d, c->nzh, nbx, \
                                    (long long)c->ny * c->nzh, c->ny, (long long)c->NK,                 \
                                    (long long)c->nzh, twp<R>(c, 0))
    FFT_DISPATCH(c->nx, CALLX);
#undef CALLX
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

template <typename R, int IN>
int fwd(evpgpu_ctx *c, const double *real_in, const float *inA, int ncomp, typename CT<R>::T *spec)
{
    // the Stockham z kernels take whole groups of TB lines
    if (!is_stockham_len(c->nz) || (c->nx * c->ny) % TB != 0) {
        const int L = c->nz, tbl = dft_lines(L);
        const long long nl = (long long)c->nx * c->ny;
        hipLaunchKernelGGL((k_dft_z_r2c<R, IN>), dim3((unsigned)((nl + tbl - 1) / tbl), ncomp), dim3(NT),
                           sizeof(typename CT<R>::T) * (2 * tbl + 1) * L, c->stream, real_in, inA, spec, L, nl, c->N,
                           c->NK, twp<R>(c, 2));
        if (hipGetLastError() != hipSuccess) return EVPGPU_ERR_HIP;
    } else {
    dim3 g((unsigned)(c->nx * c->ny / TB), ncomp);
#define CALLZ(L) hipLaunchKernelGGL((k_fft_z_r2c<R, L, IN>), g, dim3(NT), 0, c->stream, real_in, inA, \
                                    spec, c->N, c->NK, twp<R>(c, 2))
    FFT_DISPATCH(c->nz, CALLZ);
#undef CALLZ
    }
    int st = pass_y<R, false>(c, spec, ncomp);
    if (st) return st;
    return pass_x<R, false>(c, spec, ncomp);
}

template <typename R, int MODE>
int inv(evpgpu_ctx *c, typename CT<R>::T *spec, int ncomp, double *target, const double *addc,
        double scale, const float *fa = nullptr, float *fo = nullptr, const double *base = nullptr)
{
    int st = pass_x<R, true>(c, spec, ncomp);
    if (st) return st;
    st = pass_y<R, true>(c, spec, ncomp);
    if (st) return st;
    double a[6] = {0, 0, 0, 0, 0, 0};
    if (addc)
        for (int i = 0; i < ncomp && i < 6; i++) a[i] = addc[i];
    if (!is_stockham_len(c->nz) || (c->nx * c->ny) % TB != 0) {
        const int L = c->nz, tbl = dft_lines(L);
        const long long nl = (long long)c->nx * c->ny;
        hipLaunchKernelGGL((k_dft_z_c2r<R, MODE>), dim3((unsigned)((nl + tbl - 1) / tbl), ncomp), dim3(NT),
                           sizeof(typename CT<R>::T) * (2 * tbl + 1) * L, c->stream, spec, target, fa, fo, base,
                           L, nl, c->N, c->NK, a[0], a[1], a[2], a[3], a[4], a[5], scale, twp<R>(c, 2));
        return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
    }
    dim3 g((unsigned)(c->nx * c->ny / TB), ncomp);
#define CALLZI(L) hipLaunchKernelGGL((k_fft_z_c2r<R, L, MODE>), g, dim3(NT), 0, c->stream, spec, target, \
                                     fa, fo, base, c->N, c->NK, a[0], a[1], a[2], a[3], a[4], a[5],   \
                                     scale, twp<R>(c, 2))
    FFT_DISPATCH(c->nz, CALLZI);
#undef CALLZI
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

} // namespace

int fft_twiddles(evpgpu_ctx *c)
{
    int dims[3] = {c->nx, c->ny, c->nz};
    for (int a = 0; a < 3; a++) {
        int L = dims[a];
        double2 *h64 = (double2 *)malloc(sizeof(double2) * L);
        float2 *h32 = (float2 *)malloc(sizeof(float2) * L);
        for (int m = 0; m < L; m++) {
            // exp(-2 pi i m / L); the quarter-period values are exact
            double ang = 2.0 * M_PI * (double)m / (double)L;
            double cr = cos(ang), si = -sin(ang);
            if (4 * m == L) { cr = 0.0; si = -1.0; }
            if (2 * m == L) { cr = -1.0; si = 0.0; }
            if (4 * m == 3 * L) { cr = 0.0; si = 1.0; }
            h64[m] = make_double2(cr, si);
            h32[m] = make_float2((float)cr, (float)si);
        }
        EVP_CHECK(hipMalloc(&c->tw64[a], sizeof(double2) * L));
        EVP_CHECK(hipMalloc(&c->tw32[a], sizeof(float2) * L));
        EVP_CHECK(hipMemcpy(c->tw64[a], h64, sizeof(double2) * L, hipMemcpyHostToDevice));
        EVP_CHECK(hipMemcpy(c->tw32[a], h32, sizeof(float2) * L, hipMemcpyHostToDevice));
        free(h64);
        free(h32);
    }
    return EVPGPU_OK;
}

// forward transform of ncomp real FP64 fields into c->spec (precision chosen)
int fft_forward(evpgpu_ctx *c, const double *real_in, int ncomp, int precision)
{
    if (precision == EVPGPU_MIXED) return fwd<float, 0>(c, real_in, nullptr, ncomp, (float2 *)c->spec);
    return fwd<double, 0>(c, real_in, nullptr, ncomp, (double2 *)c->spec);
}

// mixed loop: FP32 transform of (sig - anchor), anchor stored in FP32
// FP64 R2C of ncomp real fields into a caller-chosen spectrum buffer
int fft_forward_to(evpgpu_ctx *c, const double *real_in, int ncomp, double2 *dst)
{
    return fwd<double, 0>(c, real_in, nullptr, ncomp, dst);
}

int fft_forward_delta(evpgpu_ctx *c, const double *sig, const float *anchor)
{
    return fwd<float, 1>(c, sig, anchor, 6, (float2 *)c->spec);
}

// mixed loop: FP64 transform of the FP32 anchor field
int fft_forward_anchor(evpgpu_ctx *c, const float *anchor)
{
    return fwd<double, 2>(c, nullptr, anchor, 6, (double2 *)c->spec);
}

// mixed loop: E += add + cA + re/N from the FP32 delta spectrum
int fft_inverse_delta(evpgpu_ctx *c, double *E, const double *add, const float *cA, double scale)
{
    return inv<float, 2>(c, (float2 *)c->spec, 6, E, add, scale, cA, nullptr, nullptr);
}

// mixed loop: cA = re/N of the FP64 anchor spectrum
int fft_inverse_anchor(evpgpu_ctx *c, float *cA, double scale)
{
    return inv<double, 3>(c, (double2 *)c->spec, 6, nullptr, nullptr, scale, nullptr, cA, nullptr);
}

// target = base + re*scale of the FP64 spectrum (6 comps)
int fft_inverse_set(evpgpu_ctx *c, double *target, const double *base, double scale)
{
    return inv<double, 4>(c, (double2 *)c->spec, 6, target, nullptr, scale, nullptr, nullptr, base);
}

// inverse of c->spec (ncomp comps) accumulated into target: t += add[c] + re*scale
int fft_inverse_update(evpgpu_ctx *c, int ncomp, int precision, double *target,
                       const double *add_const, double scale)
{
    if (precision == EVPGPU_MIXED)
        return inv<float, 0>(c, (float2 *)c->spec, ncomp, target, add_const, scale);
    return inv<double, 0>(c, (double2 *)c->spec, ncomp, target, add_const, scale);
}

// inverse of an FP64 spectrum (destroyed) into out = re*scale
int fft_inverse_real(evpgpu_ctx *c, const double2 *cin, int ncomp, double *out, double scale)
{
    return inv<double, 1>(c, (double2 *)cin, ncomp, out, nullptr, scale);
}
