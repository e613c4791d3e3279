// fft_core.h -- Stockham FFT building blocks shared by the FFT passes
// (fft.hip) and the fused plane / line kernels (fft_fused.hip): complex
// helpers, radix-2/3/4/5/8 butterflies, the radix plan of a 2^m 3^a 5^b length
// and a generic LDS stage over any set of lines.
#pragma once

#include <hip/hip_runtime.h>

namespace fftc {

template <typename R> struct CT;
template <> struct CT<double> { typedef double2 T; };
template <> struct CT<float> { typedef float2 T; };

// Move `total` elements into LDS: element e = threadIdx.x + u NTH is
// load(e) (a value of type T), then store(e, value).  A loop that stores
// each value as it arrives waits on memory once per element; here a thread's
// loads go out together, CH at a time (CH bounds the registers they hold),
// and are waited for once per chunk.
template <int NTH, int PER, int CH, typename T, typename LoadF, typename StoreF>
__device__ __forceinline__ void staged_copy(int total, LoadF load, StoreF store)
{
#pragma unroll
    for (int u0 = 0; u0 < PER; u0 += CH) {
        T v[CH];
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int e = threadIdx.x + (u0 + u) * NTH;
            v[u] = T{};
            if (u0 + u < PER && e < total) v[u] = load(e);
        }
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int e = threadIdx.x + (u0 + u) * NTH;
            if (u0 + u < PER && e < total) store(e, v[u]);
        }
    }
}

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return {a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return {a.x - b.x, a.y - b.y}; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename C> __device__ __forceinline__ C cmi(C a)
{
    if (INV) return {-a.y, a.x};
    return {a.y, -a.x};
}

template <bool INV, typename C> __device__ __forceinline__ void bf2(C &a, C &b)
{
    C t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <bool INV, typename C> __device__ __forceinline__ void bf4(C v[4])
{
    C s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    C s13 = cadd(v[1], v[3]), d13 = cmi<INV>(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
}

template <bool INV, typename C> __device__ __forceinline__ void bf8(C v[8])
{
    typedef decltype(v[0].x) R;
    const R h = (R)0.70710678118654752440;
    C e[4] = {v[0], v[2], v[4], v[6]};
    C o[4] = {v[1], v[3], v[5], v[7]};
    bf4<INV>(e);
    bf4<INV>(o);
    // o[k] *= W8^k, W8 = exp(-+ 2 pi i / 8)
    C w1 = INV ? C{h, h} : C{h, -h};
    C w3 = INV ? C{-h, h} : C{-h, -h};
    o[1] = cmul(o[1], w1);
    o[2] = cmi<INV>(o[2]);
    o[3] = cmul(o[3], w3);
    for (int k = 0; k < 4; k++) {
        v[k] = cadd(e[k], o[k]);
        v[k + 4] = csub(e[k], o[k]);
    }
}

// radix 3: y0 = a + b + c, y1,2 = a - (b + c)/2 -+ i (sqrt 3 / 2)(b - c)
// (forward; the signs of the i terms swap for the inverse)
template <bool INV, typename C> __device__ __forceinline__ void bf3(C v[3])
{
    typedef decltype(v[0].x) R;
    const R h = (R)0.86602540378443864676; // sqrt(3)/2
    C s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    C m = {v[0].x - (R)0.5 * s.x, v[0].y - (R)0.5 * s.y};
    C r = cmi<INV>(C{h * d.x, h * d.y});
    v[0] = cadd(v[0], s);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
}

// radix 5 from the real and imaginary parts of w = exp(-+2 pi i / 5):
// y_k = a + sum_j w^(jk) x_j, pairing x1 + x4, x2 + x3 (cosine terms) and
// x1 - x4, x2 - x3 (sine terms)
template <bool INV, typename C> __device__ __forceinline__ void bf5(C v[5])
{
    typedef decltype(v[0].x) R;
    const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410; // cos 2pi/5, cos 4pi/5
    const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917; // sin 2pi/5, sin 4pi/5
    C p1 = cadd(v[1], v[4]), m1 = csub(v[1], v[4]);
    C p2 = cadd(v[2], v[3]), m2 = csub(v[2], v[3]);
    C a = v[0];
    C t1 = {a.x + c1 * p1.x + c2 * p2.x, a.y + c1 * p1.y + c2 * p2.y};
    C t2 = {a.x + c2 * p1.x + c1 * p2.x, a.y + c2 * p1.y + c1 * p2.y};
    // -i (s1 m1 + s2 m2) and -i (s2 m1 - s1 m2) forward, +i inverse
    C u1 = cmi<INV>(C{s1 * m1.x + s2 * m2.x, s1 * m1.y + s2 * m2.y});
    C u2 = cmi<INV>(C{s2 * m1.x - s1 * m2.x, s2 * m1.y - s1 * m2.y});
    v[0] = cadd(a, cadd(p1, p2));
    v[1] = cadd(t1, u1);
    v[4] = csub(t1, u1);
    v[2] = cadd(t2, u2);
    v[3] = csub(t2, u2);
}

template <bool INV, int RAD, typename C> __device__ __forceinline__ void bfly(C v[RAD])
{
    if constexpr (RAD == 2) bf2<INV>(v[0], v[1]);
    else if constexpr (RAD == 3) bf3<INV>(v);
    else if constexpr (RAD == 4) bf4<INV>(v);
    else if constexpr (RAD == 5) bf5<INV>(v);
    else bf8<INV>(v);
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }
// multiplicity of the factor f in L
constexpr int fcount(int L, int f) { return (L % f) ? 0 : 1 + fcount(L / f, f); }
constexpr int ipowi(int b, int e) { return e == 0 ? 1 : b * ipowi(b, e - 1); }
// radix plan of L = 2^m 3^a 5^b: the radix-5 stages, the radix-3 stages,
// then radix 8 and the 4 or 2 left over
constexpr int radix_at(int L, int s)
{
    const int c5 = fcount(L, 5), c3 = fcount(L, 3);
    if (s < c5) return 5;
    if (s < c5 + c3) return 3;
    s -= c5 + c3;
    int m = ilog2(L / ipowi(5, c5) / ipowi(3, c3)), q = m / 3, r = m % 3;
    return s < q ? 8 : (r == 2 ? 4 : 2);
}
constexpr int nstages(int L)
{
    const int c5 = fcount(L, 5), c3 = fcount(L, 3);
    int m = ilog2(L / ipowi(5, c5) / ipowi(3, c3));
    return c5 + c3 + m / 3 + (m % 3 ? 1 : 0);
}

// One Stockham stage of length-L FFTs on NL lines held in LDS, element e of
// line t at lds[t*LS + e*ES].  MAP 0 gives consecutive threads consecutive
// butterflies of one line (lines along the fast LDS axis), MAP 1 consecutive
// lines (lines across it), so that LDS accesses stay bank-friendly either
// way.  Twiddles: a table of exp(-2 pi i m / (L*TWS)), read at stride TWS.
template <bool INV, int L, int RAD, int NS, int NL, int LS, int ES, int MAP, int NTH, int TWS, typename C>
__device__ __forceinline__ void gstage(C *lds, const C *tw)
{
    constexpr int NB = L / RAD;
    constexpr int TOT = NB * NL;
    constexpr int PER = (TOT + NTH - 1) / NTH;
    C v[PER][RAD];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        int b = threadIdx.x + q * NTH;
        if (b < TOT) {
            int t = MAP ? b % NL : b / NB, j = MAP ? b / NL : b % NB;
#pragma unroll
            for (int r = 0; r < RAD; r++) v[q][r] = lds[t * LS + (j + r * NB) * ES];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++) {
        int b = threadIdx.x + q * NTH;
        if (b < TOT) {
            int t = MAP ? b % NL : b / NB, j = MAP ? b / NL : b % NB;
            int k = j % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < RAD; r++) {
                    C w = tw[((r * k * (L / (NS * RAD))) % L) * TWS];
                    if (INV) w.y = -w.y;
                    v[q][r] = cmul(v[q][r], w);
                }
            }
            bfly<INV, RAD>(v[q]);
            int d = (j / NS) * NS * RAD + k;
#pragma unroll
            for (int r = 0; r < RAD; r++) lds[t * LS + (d + r * NS) * ES] = v[q][r];
        }
    }
    __syncthreads();
}

template <bool INV, int L, int S, int NS, int NL, int LS, int ES, int MAP, int NTH, int TWS, typename C>
__device__ __forceinline__ void gstages(C *lds, const C *tw)
{
    if constexpr (S < nstages(L)) {
        constexpr int RAD = radix_at(L, S);
        gstage<INV, L, RAD, NS, NL, LS, ES, MAP, NTH, TWS>(lds, tw);
        gstages<INV, L, S + 1, NS * RAD, NL, LS, ES, MAP, NTH, TWS>(lds, tw);
    }
}

// Real transforms of length 2M through M-point complex ones: the reals
// (x_2n, x_2n+1) are packed as Z_n = x_2n + i x_2n+1; twz is the table of
// length 2M (exp(-2 pi i m / 2M)).
// z line of M packed complex values Z -> the M+1 half-spectrum values X of
// the 2M reals (in place; thread pairs k and M-k)
template <typename R, int M, typename C>
__device__ __forceinline__ void split_forward(C *row, int k, const C *twz)
{
    if (k == 0) {
        C z = row[0];
        row[0] = C{z.x + z.y, (R)0};
        row[M] = C{z.x - z.y, (R)0};
        return;
    }
    C a = row[k], b = row[M - k];
    // Fe = (Z_k + conj(Z_{M-k}))/2, Fo = (Z_k - conj(Z_{M-k}))/(2i)
    C fe = {(a.x + b.x) * (R)0.5, (a.y - b.y) * (R)0.5};
    C fo = {(a.y + b.y) * (R)0.5, -(a.x - b.x) * (R)0.5};
    C w = twz[k]; // exp(-2 pi i k / 2M)
    C wf = cmul(w, fo);
    row[k] = cadd(fe, wf);
    if (2 * k != M) row[M - k] = C{fe.x - wf.x, -(fe.y - wf.y)}; // conj(Fe - W^k Fo)
}

// the inverse: half spectrum X_0..X_M -> M packed complex values Z whose
// inverse FFT gives (x_2n, x_2n+1) of the C2R of the Hermitian extension
template <typename R, int M, typename C>
__device__ __forceinline__ void split_inverse(C *row, int k, const C *twz)
{
    if (k == 0) {
        R a = row[0].x, b = row[M].x; // imaginary parts dropped
        row[0] = C{a + b, a - b};
        return;
    }
    C xk = row[k], xm = row[M - k];
    C w = twz[k];
    w.y = -w.y; // exp(+2 pi i k / 2M)
    // Z_k = (X_k + conj(X_{M-k})) + i w^k (X_k - conj(X_{M-k}))
    C s = {xk.x + xm.x, xk.y - xm.y}, d = {xk.x - xm.x, xk.y + xm.y};
    C wd = cmul(w, d);
    C zk = {s.x - wd.y, s.y + wd.x};
    if (2 * k != M) {
        // Z_{M-k} = (X_{M-k} + conj(X_k)) + i w^{M-k} (X_{M-k} - conj(X_k)), w^{M-k} = -conj(w^k)
        C s2 = {xm.x + xk.x, xm.y - xk.y}, d2 = {xm.x - xk.x, xm.y + xk.y};
        C w2 = {-w.x, w.y};
        C wd2 = cmul(w2, d2);
        row[M - k] = C{s2.x - wd2.y, s2.y + wd2.x};
    }
    row[k] = zk;
}

} // namespace fftc
