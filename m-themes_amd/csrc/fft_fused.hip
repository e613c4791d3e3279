// fft_fused.hip -- one Green iteration's transforms in three passes over the
// spectrum instead of seven (fft.hip's z, y, x passes each way plus the Green
// pass), for grids whose (y, z) half-spectrum plane fits in LDS.
//
//   k_zy_r2c   one workgroup per (component, x): the ny x nz real plane is
//              read once, transformed along z as ny real FFTs (each as a
//              length-nz/2 complex FFT of the even/odd samples plus the
//              split step), then along y for every kz, and written once as
//              the ny x (nz/2+1) half-spectrum plane;
//   k_x_green  one workgroup per run of TBX (y, kz) columns: all 6 components
//              of those lines are transformed along x, the Green operator is
//              applied at each wave vector (green_point.h, with the
//              antisymmetric accumulator update at the split wave vectors),
//              and transformed back;
//   k_yz_c2r   the mirror of k_zy_r2c, ending in the strain update.
//
// The results equal fft.hip's up to rounding: the same half-spectrum
// convention, the real transforms built from half-length complex ones, and
// the C2R taking the real part of the Hermitian extension (the imaginary
// parts of the kz = 0 and kz = nz/2 entries are dropped as fft.hip drops
// them).  LDS per workgroup: ny (nz/2+1) complex numbers for the plane
// kernels (133 KB at 128^2 in FP64), 6 TBX (nx+1) for the line kernel.
#include <math.h>

#include "evpgpu_impl.h"
#include "fft_core.h"
#include "green_point.h"

namespace {

using namespace fftc;

constexpr int NTH = 1024;        // threads per workgroup of the plane kernels
constexpr int NTX = 256;         // of the x-line kernel (up to 3 workgroups per CU)
constexpr int LDS_MAX = 160 * 1024;

template <typename R, int NY, int NZ> struct Plane {
    typedef typename CT<R>::T C;
    static constexpr int M = NZ / 2;      // half length of a z line
    static constexpr int P = M + 1;       // row pitch = nz/2 + 1 (half spectrum)
    static constexpr int BYTES = NY * P * (int)sizeof(C);
};

// IN 0: FP64 field; IN 1: FP64 field minus FP32 anchor; IN 2: FP32 field
template <typename R, int NY, int NZ, int IN>
__global__ __launch_bounds__(NTH) void k_zy_r2c(const double *__restrict__ in, const float *__restrict__ inA,
                                                typename CT<R>::T *__restrict__ out, long long N, long long NK,
                                                int nx, const typename CT<R>::T *__restrict__ twy,
                                                const typename CT<R>::T *__restrict__ twz,
                                                const int *__restrict__ skip)
{
    typedef Plane<R, NY, NZ> PL;
    typedef typename PL::C C;
    constexpr int M = PL::M, P = PL::P;
    if (skip && *skip) return;
    __shared__ C lds[NY * P];
    // the twiddle tables in LDS, loaded with the plane: read from memory at
    // every stage, each stage waited on them
    __shared__ C s_twy[NY], s_twz[NZ];
    if (threadIdx.x < NY) s_twy[threadIdx.x] = twy[threadIdx.x];
    else if (threadIdx.x < NY + NZ) s_twz[threadIdx.x - NY] = twz[threadIdx.x - NY];
    const int x = blockIdx.x, comp = blockIdx.y;
    const long long base = comp * N + (long long)x * NY * NZ;
    // every load of the plane goes out before the first wait (a loop that
    // stores each value as it arrives waits once per element)
    constexpr int NIN = NY * M, PER = (NIN + NTH - 1) / NTH;
    double2 v[PER];
    float2 w[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int i = threadIdx.x + u * NTH;
        v[u] = make_double2(0.0, 0.0);
        w[u] = make_float2(0.f, 0.f);
        if (NIN % NTH != 0 && i >= NIN) continue;
        if (IN != 2) v[u] = reinterpret_cast<const double2 *>(in + base)[i];
        if (IN != 0) w[u] = reinterpret_cast<const float2 *>(inA + base)[i];
    }
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int i = threadIdx.x + u * NTH;
        if (NIN % NTH != 0 && i >= NIN) continue;
        const int y = i / M, n = i % M;
        R a, b;
        if (IN == 0) {
            a = (R)v[u].x;
            b = (R)v[u].y;
        } else if (IN == 1) {
            a = (R)(v[u].x - (double)w[u].x);
            b = (R)(v[u].y - (double)w[u].y);
        } else {
            a = (R)w[u].x;
            b = (R)w[u].y;
        }
        lds[y * P + n] = C{a, b};
    }
    __syncthreads();
    // z: NY rows of M packed values (twiddles of length M = the nz table at stride 2)
    gstages<false, M, 0, 1, NY, P, 1, 0, NTH, 2>(lds, (const C *)s_twz);
    for (int i = threadIdx.x; i < NY * (M / 2 + 1); i += NTH) {
        const int y = i / (M / 2 + 1), k = i % (M / 2 + 1);
        split_forward<R, M>(lds + y * P, k, (const C *)s_twz);
    }
    __syncthreads();
    // y: P columns of length NY
    gstages<false, NY, 0, 1, P, 1, P, 1, NTH, 1>(lds, (const C *)s_twy);
    C *dst = out + comp * NK + (long long)x * NY * P;
    for (int i = threadIdx.x; i < NY * P; i += NTH) dst[i] = lds[i];
}

// MODE 0: target += add + re*scale; 1: target = re*scale; 2: target += add +
// fa + re*scale; 3: fo = (float)(re*scale); 4: target = base + re*scale.
// add is addd[comp] when addd is given (device memory), else a0..a5.
template <typename R, int NY, int NZ, int MODE>
__global__ __launch_bounds__(NTH) void k_yz_c2r(const typename CT<R>::T *__restrict__ in, double *target,
                                                const float *__restrict__ fa, float *__restrict__ fo,
                                                const double *__restrict__ basef, long long N, long long NK,
                                                double a0, double a1, double a2, double a3, double a4,
                                                double a5, const double *__restrict__ addd, double scale,
                                                const typename CT<R>::T *__restrict__ twy,
                                                const typename CT<R>::T *__restrict__ twz,
                                                const int *__restrict__ skip)
{
    typedef Plane<R, NY, NZ> PL;
    typedef typename PL::C C;
    constexpr int M = PL::M, P = PL::P;
    if (skip && *skip) return;
    __shared__ C lds[NY * P];
    // the twiddle tables in LDS, loaded with the spectrum plane
    __shared__ C s_twy[NY], s_twz[NZ];
    if (threadIdx.x < NY) s_twy[threadIdx.x] = twy[threadIdx.x];
    else if (threadIdx.x < NY + NZ) s_twz[threadIdx.x - NY] = twz[threadIdx.x - NY];
    const int x = blockIdx.x, comp = blockIdx.y;
    const C *src = in + comp * NK + (long long)x * NY * P;
    {
        // the loads first (in chunks of CH, which bounds the registers they
        // hold), then the LDS stores
        constexpr int NIN = NY * P, PER = (NIN + NTH - 1) / NTH, CH = sizeof(R) == 8 ? 4 : 16;
#pragma unroll
        for (int u0 = 0; u0 < PER; u0 += CH) {
            C v[CH];
#pragma unroll
            for (int u = 0; u < CH; u++) {
                const int i = threadIdx.x + (u0 + u) * NTH;
                v[u] = C{(R)0, (R)0};
                if (u0 + u < PER && (NIN % NTH == 0 || i < NIN)) v[u] = src[i];
            }
#pragma unroll
            for (int u = 0; u < CH; u++) {
                const int i = threadIdx.x + (u0 + u) * NTH;
                if (u0 + u < PER && (NIN % NTH == 0 || i < NIN)) lds[i] = v[u];
            }
        }
    }
    __syncthreads();
    gstages<true, NY, 0, 1, P, 1, P, 1, NTH, 1>(lds, (const C *)s_twy);
    for (int i = threadIdx.x; i < NY * (M / 2 + 1); i += NTH) {
        const int y = i / (M / 2 + 1), k = i % (M / 2 + 1);
        split_inverse<R, M>(lds + y * P, k, (const C *)s_twz);
    }
    __syncthreads();
    gstages<true, M, 0, 1, NY, P, 1, 0, NTH, 2>(lds, (const C *)s_twz);
    const double add = addd ? addd[comp]
                            : comp == 0 ? a0 : comp == 1 ? a1 : comp == 2 ? a2 : comp == 3 ? a3 : comp == 4 ? a4 : a5;
    const long long base = comp * N + (long long)x * NY * NZ;
    // (loading the read-modify-write operands of all the thread's elements
    // before the first wait was measured slower here: 70 us against 63 us)
    for (int i = threadIdx.x; i < NY * M; i += NTH) {
        const int y = i / M, n = i % M;
        C z = lds[y * P + n];
        double r0 = (double)z.x * scale, r1 = (double)z.y * scale;
        double2 *t2 = reinterpret_cast<double2 *>(target + base);
        if (MODE == 0) {
            double2 t = t2[i];
            t2[i] = make_double2(t.x + (add + r0), t.y + (add + r1));
        } else if (MODE == 1) {
            t2[i] = make_double2(r0, r1);
        } else if (MODE == 2) {
            double2 t = t2[i];
            float2 f = reinterpret_cast<const float2 *>(fa + base)[i];
            t2[i] = make_double2(t.x + (add + (double)f.x + r0), t.y + (add + (double)f.y + r1));
        } else if (MODE == 3) {
            reinterpret_cast<float2 *>(fo + base)[i] = make_float2((float)r0, (float)r1);
        } else {
            double2 b = reinterpret_cast<const double2 *>(basef + base)[i];
            t2[i] = make_double2(b.x + r0, b.y + r1);
        }
    }
}

// LDS pitch of an x line (complex values).  The tile is written and read
// column-fastest: lane l takes column l % TBX of x = l / TBX.  The LDS serves
// 32 lanes of 8-byte (FP32 complex) accesses at once, and element j * PX + x
// lands in bank pair (j PX + x) mod 32: distinct for every lane of the 32
// when PX = 32 / TBX (mod 32), the smallest such pitch >= NX.  FP64 keeps
// NX + 1 (the same rule for its 16-byte accesses measured no faster).
template <typename R> constexpr int px(int nx, int tbx)
{
    if (sizeof(R) != 4) return nx + 1;
    const int want = (32 / tbx) % 32;
    return nx + ((want - nx % 32) % 32 + 32) % 32;
}

// GMODE 0: acc += antisym; 2: acc += antisym + skA; 3: skA = antisym
template <typename R, int NX, int TBX, int GMODE>
__global__ __launch_bounds__(NTX) void k_x_green(typename CT<R>::T *__restrict__ spec, double2 *__restrict__ acc,
                                                 float2 *__restrict__ skA, GreenParams G, int ny, int nz,
                                                 int nzh, long long NK, const typename CT<R>::T *__restrict__ twx,
                                                 const int *__restrict__ skip)
{
    typedef typename CT<R>::T C;
    constexpr int PX = px<R>(NX, TBX);
    constexpr int NL = 6 * TBX;
    if (skip && *skip) return;
    __shared__ C lds[NL * PX];
    __shared__ R c0s[36];
    // the twiddle table in LDS, loaded with the tile
    __shared__ C s_tw[NX];
    for (int i = threadIdx.x; i < NX; i += NTX) s_tw[i] = twx[i];
    if (threadIdx.x < 36) c0s[threadIdx.x] = (R)G.ga[threadIdx.x];
    // The x lines are the columns q = y*nzh + kz of the [x][y][kz] half
    // spectrum; a workgroup takes TBX consecutive columns q0.. (runs that
    // start on 128-byte boundaries, as ny*nzh is a multiple of TBX).
    // Line (c, j) = component c, column q0 + j; element x at spec[c*NK + x*ncol + q0 + j].
    const int ncol = ny * nzh;
    // Workgroups go to the 8 XCDs round robin (block b to XCD b mod 8), and
    // each XCD has its own L2.  A tile row is 64 B when TBX complex values
    // fill half a 128-byte line, so the two tiles that share the lines are
    // given to blocks b and b + 8 (the same XCD, dispatched together):
    // within each group of 16 blocks, block g*16 + r takes tile
    // g*16 + 2 (r mod 8) + r / 8.
    int tile = blockIdx.x;
    if (blockIdx.x < gridDim.x / 16 * 16) {
        const int g = blockIdx.x / 16, r = blockIdx.x % 16;
        tile = g * 16 + (r % 8) * 2 + r / 8;
    }
    const int q0 = tile * TBX;
    const int nk = min(TBX, ncol - q0);
    {
        // all loads of the tile first, then the transposing LDS stores
        constexpr int NIN = NL * NX, PER = (NIN + NTX - 1) / NTX;
        C v[PER];
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int i = threadIdx.x + u * NTX;
            const int c = i / (TBX * NX), r = i % (TBX * NX), x = r / TBX, j = r % TBX;
            v[u] = C{(R)0, (R)0};
            if ((NIN % NTX == 0 || i < NIN) && j < nk) v[u] = spec[c * NK + (long long)x * ncol + q0 + j];
        }
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int i = threadIdx.x + u * NTX;
            const int c = i / (TBX * NX), r = i % (TBX * NX), x = r / TBX, j = r % TBX;
            if (NIN % NTX == 0 || i < NIN) lds[(c * TBX + j) * PX + x] = v[u];
        }
    }
    __syncthreads();
    gstages<false, NX, 0, 1, NL, PX, 1, 0, NTX, 1>(lds, (const C *)s_tw);
    // Green result at (x, column q0 + j) back into LDS, antisymmetric part
    // into the accumulator
    auto put = [&](int x, int j, const C e6[6], const double wr[3], const double wi[3], bool split) {
#pragma unroll
        for (int c = 0; c < 6; c++) lds[(c * TBX + j) * PX + x] = e6[c];
        // antisymmetric parts are kept at the split wave vectors only (the
        // others follow from the increment's strain change at its end)
        if (!split) return;
        const long long id = (long long)x * ncol + q0 + j;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            if (GMODE == 3) {
                skA[c * NK + id] = make_float2((float)wr[c], (float)wi[c]);
            } else {
                double2 w = acc[c * NK + id];
                double ax = wr[c], ay = wi[c];
                if (GMODE == 2) {
                    float2 k = skA[c * NK + id];
                    ax += (double)k.x;
                    ay += (double)k.y;
                }
                acc[c * NK + id] = make_double2(w.x + ax, w.y + ay);
            }
        }
    };
#pragma unroll 1
    for (int i = threadIdx.x; i < TBX * NX; i += NTX) {
        // keep the 36 Green coefficients in LDS: hoisted out of this loop
        // they would occupy 72 registers for the whole loop
        __asm__ volatile("" ::: "memory");
        // consecutive lanes take consecutive x of one column: conflict-free
        // LDS rows, and a wave-uniform (y, kz)
        const int x = i % NX, j = i / NX;
        const int q = q0 + j;
        if (j >= nk || (q == 0 && x == 0)) continue; // k = 0 below
        C s6[6], e6[6];
        double wr[3], wi[3];
#pragma unroll
        for (int c = 0; c < 6; c++) s6[c] = lds[(c * TBX + j) * PX + x];
        const int y = q / nzh, kz = q - y * nzh;
        bool split;
        gop::green_point<R>(c0s, x, y, kz, NX, ny, nz, s6, e6, wr, wi, split);
        put(x, j, e6, wr, wi, split);
    }
    if (q0 == 0 && threadIdx.x == 0) {
        C s6[6], e6[6];
        double wr[3], wi[3];
#pragma unroll
        for (int c = 0; c < 6; c++) s6[c] = lds[c * TBX * PX];
        gop::green_zero<R>(c0s, G.s0, G.elast_bc, NX, ny, nz, s6, e6, wr, wi);
        put(0, 0, e6, wr, wi, false);
    }
    __syncthreads();
    gstages<true, NX, 0, 1, NL, PX, 1, 0, NTX, 1>(lds, (const C *)s_tw);
    for (int i = threadIdx.x; i < NL * NX; i += NTX) {
        const int c = i / (TBX * NX), r = i % (TBX * NX), x = r / TBX, j = r % TBX;
        if (j < nk) spec[c * NK + (long long)x * ncol + q0 + j] = lds[(c * TBX + j) * PX + x];
    }
}

template <typename R> const typename CT<R>::T *twp(evpgpu_ctx *c, int axis);

template <> const double2 *twp<double>(evpgpu_ctx *c, int axis) { return c->tw64[axis]; }
template <> const float2 *twp<float>(evpgpu_ctx *c, int axis) { return c->tw32[axis]; }

// lines per x-green workgroup: 6 components x TBX lines x px() in LDS,
// sized so that three workgroups share a CU's LDS (their load, transform and
// Green phases then overlap)
template <typename R, int NX> constexpr int tbx()
{
    constexpr int cs = (int)sizeof(typename CT<R>::T);
    constexpr int cap = LDS_MAX / 3 - 1024 - NX * cs; // less the twiddle table
    return 6 * 16 * px<R>(NX, 16) * cs <= cap  ? 16
           : 6 * 8 * px<R>(NX, 8) * cs <= cap  ? 8
           : 6 * 4 * px<R>(NX, 4) * cs <= cap  ? 4
                                               : 2;
}

// plane (ny = nz) and line (nx) lengths of the fused kernels: powers of two
// and the common mixed-radix lengths
#define PLANE_DISPATCH(CALL)                                   \
    if (c->ny != c->nz) return EVPGPU_ERR_SIZE;                \
    switch (c->ny) {                                           \
    case 8: CALL(8, 8); break;                                 \
    case 16: CALL(16, 16); break;                              \
    case 24: CALL(24, 24); break;                              \
    case 32: CALL(32, 32); break;                              \
    case 48: CALL(48, 48); break;                              \
    case 64: CALL(64, 64); break;                              \
    case 96: CALL(96, 96); break;                              \
    case 120: CALL(120, 120); break;                           \
    case 128: CALL(128, 128); break;                           \
    default: return EVPGPU_ERR_SIZE;                           \
    }

#define LINE_DISPATCH(CALL)             \
    switch (c->nx) {                    \
    case 8: CALL(8); break;             \
    case 16: CALL(16); break;           \
    case 24: CALL(24); break;           \
    case 32: CALL(32); break;           \
    case 48: CALL(48); break;           \
    case 64: CALL(64); break;           \
    case 96: CALL(96); break;           \
    case 120: CALL(120); break;         \
    case 128: CALL(128); break;         \
    case 192: CALL(192); break;         \
    case 240: CALL(240); break;         \
    case 256: CALL(256); break;         \
    default: return EVPGPU_ERR_SIZE;    \
    }

template <typename R, int IN>
int zy(evpgpu_ctx *c, const double *in, const float *inA)
{
    typedef typename CT<R>::T C;
    C *spec = (C *)c->spec;
#define CALLP(NY_, NZ_)                                                                                   \
    hipLaunchKernelGGL((k_zy_r2c<R, NY_, NZ_, IN>), dim3(c->nx, 6), dim3(NTH), 0, c->stream, in, inA, spec, \
                       c->N, c->NK, c->nx, twp<R>(c, 1), twp<R>(c, 2), c->skip)
    PLANE_DISPATCH(CALLP);
#undef CALLP
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

// addc: host values of the constant added per component (or null: none);
// addd: the same in device memory (read by the kernel, overrides addc)
template <typename R, int MODE>
int yz(evpgpu_ctx *c, double *target, const double *addc, double scale, const float *fa, float *fo,
       const double *basef, const double *addd = nullptr)
{
    typedef typename CT<R>::T C;
    const C *spec = (const C *)c->spec;
    double a[6] = {0, 0, 0, 0, 0, 0};
    if (addc)
        for (int i = 0; i < 6; i++) a[i] = addc[i];
#define CALLP(NY_, NZ_)                                                                                    \
    hipLaunchKernelGGL((k_yz_c2r<R, NY_, NZ_, MODE>), dim3(c->nx, 6), dim3(NTH), 0, c->stream, spec, target, \
                       fa, fo, basef, c->N, c->NK, a[0], a[1], a[2], a[3], a[4], a[5], addd, scale,          \
                       twp<R>(c, 1), twp<R>(c, 2), c->skip)
    PLANE_DISPATCH(CALLP);
#undef CALLP
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

template <typename R, int GMODE>
int xg(evpgpu_ctx *c)
{
    typedef typename CT<R>::T C;
    const GreenParams G = green_params(c);
#define CALLL(NX_)                                                                                        \
    {                                                                                                    \
        constexpr int T = tbx<R, NX_>();                                                                 \
        hipLaunchKernelGGL((k_x_green<R, NX_, T, GMODE>), dim3((c->ny * c->nzh + T - 1) / T), dim3(NTX), 0,  \
                           c->stream, (C *)c->spec, c->acc, c->skA, G, c->ny, c->nz, c->nzh, c->NK,      \
                           twp<R>(c, 0), c->skip);                                                       \
    }
    LINE_DISPATCH(CALLL);
#undef CALLL
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

} // namespace

// the fused path covers the plane lengths ny = nz of PLANE_DISPATCH (<= 128:
// the plane fits in LDS in FP64) and the line lengths nx of LINE_DISPATCH
bool fused_ok(const evpgpu_ctx *c)
{
    static const int planes[] = {8, 16, 24, 32, 48, 64, 96, 120, 128};
    static const int lines[] = {8, 16, 24, 32, 48, 64, 96, 120, 128, 192, 240, 256};
    bool p = false, l = false;
    for (int n : planes) p = p || (c->ny == n && c->nz == n);
    for (int n : lines) l = l || c->nx == n;
    return p && l;
}

// FP64 iteration: spec = R2C(sigma) ; Green (acc += antisym) ; E += add + C2R/N
// (add: 6 values in device memory)
int fused_iteration_fp64(evpgpu_ctx *c, const double *sig, double *E, const double *add, double scale)
{
    tmark(c, 1, 0);
    int st = zy<double, 0>(c, sig, nullptr);
    tmark(c, 1, 1);
    tmark(c, 2, 0);
    if (!st) st = xg<double, 0>(c);
    tmark(c, 2, 1);
    tmark(c, 3, 0);
    if (!st) st = yz<double, 0>(c, E, nullptr, scale, nullptr, nullptr, nullptr, add);
    tmark(c, 3, 1);
    return st;
}

// mixed loop, delta step (FP32): spec = R2C(sigma - sigma_A) ; Green (acc += antisym + skA) ;
// E += add + c_A + C2R/N (add: 6 values in device memory)
int fused_iteration_delta(evpgpu_ctx *c, const double *sig, const float *sigA, double *E, const double *add,
                          const float *cA, double scale)
{
    tmark(c, 1, 0);
    int st = zy<float, 1>(c, sig, sigA);
    tmark(c, 1, 1);
    tmark(c, 2, 0);
    if (!st) st = xg<float, 2>(c);
    tmark(c, 2, 1);
    tmark(c, 3, 0);
    if (!st) st = yz<float, 2>(c, E, nullptr, scale, cA, nullptr, nullptr, add);
    tmark(c, 3, 1);
    return st;
}

// mixed loop, anchor (FP64): skA = antisym(Gamma sigma_A) ; c_A = C2R(Gamma sigma_A)/N
int fused_anchor(evpgpu_ctx *c, const float *sigA, float *cA, double scale)
{
    int st = zy<double, 2>(c, nullptr, sigA);
    if (!st) st = xg<double, 3>(c);
    if (!st) st = yz<double, 3>(c, nullptr, nullptr, scale, nullptr, cA, nullptr);
    return st;
}
