// adjoint.hip -- reverse pass of the loading increments: d(loss)/d(tau0,
// tau1, theta0, theta1, n) for a loss of the macroscopic stress curve.
//
// What is differentiated.  At convergence an increment's Green loop
// (src/evolution.c:1343-1463) satisfies, per voxel, the constitutive law of
// UpdateStress with lambda = sigma,
//     S sigma + eps_p + dt edot(sigma; P(T), crss, n) = E,
// with E = Ebar + e~, e~ compatible (the span of the Green operator's
// output) and sigma in equilibrium (Gamma_eff sigma = 0 for k != 0), the
// strain-controlled components of Ebar prescribed and the stress-controlled
// components of <sigma> equal to the applied stress.  A perturbation dr of
// the right-hand side (eps_p, or edot through T, crss, n) moves the solution
// by dE = A^-1 P C_t dr, dsigma = C_t (dE - dr), where C_t = (S + dt
// d edot/d sigma)^-1 is the per-voxel tangent (symmetric positive
// definite), P the orthogonal projection onto admissible strains
// (compatible fluctuations + the free macroscopic components) and
// A = P C_t P.  For adjoint seeds (sigma_bar, e_bar) the reverse step is
//     lambda = A^-1 P (C_t sigma_bar + e_bar),  r_bar = C_t (lambda - sigma_bar),
// one conjugate-gradient solve per increment (A is self-adjoint and positive
// definite on admissible fields): the tolerance-terminated forward
// iteration itself is never differentiated.
//
// Between increments: Eps += edot dt, update_orient (which reads the
// antisymmetric displacement-gradient increment dW), harden (Voce), all
// differentiated per voxel (adjoint_voxel.h).  dW is a linear function of
// the compatible strain change: at a wave vector with a single direction n,
// e = sym(a x n) fixes a = 2 e.n - (n.e.n) n and dW = skew(a x n); at the
// Nyquist entries where Gamma_eff mixes n1 = n(k) and n2 = n(-k), the
// Green output is -1/2 sum_i (Omega_i sigma n_i) x n_i, whose symmetric part
// determines sigma up to m x m (m = n1 x n2), which the skew part does not
// see, so dW = K G^+ e with G, K the symmetric and skew parts of that map.
//
// Layout: every 6-vector of the reverse pass is stored in the orthonormal
// B-basis (the Newton solve's basis), component-major like the state.
#include <math.h>
#include <stdio.h>
#include <time.h>

#include "adjoint_voxel.h"
#include "evpgpu_impl.h"
#include "spectral.h"

struct evpgpu_adj {
    int n, ck, ncp;
    int *iters, *fp64, *f_orient, *f_harden; // per recorded increment
    evpgpu_phase ph[2];
    int nph;
    evpgpu_loading ld;
    // checkpoints: ncp segment starts + the final state, STATE_DBL * N each
    double *cp;
    evpgpu_macro *cp_m;
    double *cp_dga;
    // segment tape: ck slots of TAPE_DBL * N
    double *tape;
    // adjoint carries and work fields
    double *ep_b, *T_b, *crss_b, *G_b, *dWb, *dWb_next;
    double *sig_b, *Ct, *x, *r, *p, *t, *q;
    double *scal;   // device CG scalars
    double *h_scal; // pinned copy
    int *d_bad;     // count of voxels whose tangent is not positive definite
};

namespace {

constexpr int STATE_DBL = 73; // sig 6, E 6, W 3, epsp 6, edot 6, tran 9, crss 24, gamacum 1, gamdot 12
constexpr int TAPE_DBL = 43;  // sig 6, tran 9, crss 24, gamacum 1, dW 3
constexpr int AB = 128;       // threads per block of the per-voxel kernels
constexpr int NPAR = 2 * EVPGPU_NPARAM;

double now_s()
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

struct AdjParams {
    double dt;
    double udot_a[9];
    int orient, harden;
    int nsys[2];
    int is_gas[2];
    double nrs[2][ADJ_NS], gam0[2][ADJ_NS];
    double hard[2][ADJ_NS][ADJ_NS];
    double tau[2][ADJ_NS][3];
    double thet[2][ADJ_NS][2];
    double bn[2][ADJ_NS][3], bb[2][ADJ_NS][3];
    double seed[6]; // dL/d sigma (B-basis) per voxel
};

__device__ __forceinline__ double block_sum_write(double v, double *partials, int slot, int nval)
{
    __shared__ double ws[AB / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) ws[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < AB / 64; i++) s += ws[i];
        partials[(long long)blockIdx.x * nval + slot] = s;
    }
    return v;
}

__device__ void load_phase(const AdjParams &P, int p, adj::HardParams &H, adj::OrientParams &O)
{
    const int ns = P.nsys[p];
    H.ns = ns;
    H.dt = P.dt;
    O.ns = ns;
    O.dt = P.dt;
    for (int i = 0; i < 9; i++) O.udot_a[i] = P.udot_a[i];
    for (int s = 0; s < ns; s++) {
        for (int j = 0; j < ns; j++) H.hard[s][j] = P.hard[p][s][j];
        for (int k = 0; k < 3; k++) {
            H.tau[s][k] = P.tau[p][s][k];
            O.bn[s][k] = P.bn[p][s][k];
            O.bb[s][k] = P.bb[p][s][k];
        }
        H.thet[s][0] = P.thet[p][s][0];
        H.thet[s][1] = P.thet[p][s][1];
    }
}

// Reverse of the increment's epilogue and of the direct (fixed-sigma)
// dependence of edot and gamdot on (T, crss, n); forms the tangent and the
// right-hand side C_t sigma_bar of the adjoint solve.
__global__ __launch_bounds__(AB) void k_adj_pre(AdjParams P, long long N, const int *__restrict__ phase,
                                                const int *__restrict__ orient_idx,
                                                const double *__restrict__ sgr,
                                                const double *__restrict__ tape, double *__restrict__ ep_b,
                                                double *__restrict__ T_b, double *__restrict__ crss_b,
                                                double *__restrict__ G_b, double *__restrict__ sig_b,
                                                double *__restrict__ Ct, double *__restrict__ rhs,
                                                double *__restrict__ dW_b, double *__restrict__ partials,
                                                int *__restrict__ bad)
{
    const long long v = (long long)blockIdx.x * AB + threadIdx.x;
    double par[NPAR];
    for (int i = 0; i < NPAR; i++) par[i] = 0.0;
    if (v < N) {
        int p = phase[v] - 1;
        p = p < 0 ? 0 : (p > 1 ? 1 : p);
        const double *tsig = tape, *ttran = tape + 6 * N, *tcrss = tape + 15 * N, *tG = tape + 39 * N,
                     *tdW = tape + 40 * N;
        if (P.is_gas[p]) {
            for (int c = 0; c < 6; c++) {
                sig_b[c * N + v] = 0.0;
                rhs[c * N + v] = 0.0;
            }
            for (int c = 0; c < 21; c++) Ct[c * N + v] = 0.0;
            for (int c = 0; c < 3; c++) dW_b[c * N + v] = 0.0;
        } else {
            const int ns = P.nsys[p];
            double T[3][3], Tb[3][3], crss[2 * ADJ_NS], crssb[2 * ADJ_NS], t6[6], sigB[6], dW[3];
            for (int i = 0; i < 9; i++) {
                T[i / 3][i % 3] = ttran[i * N + v];
                Tb[i / 3][i % 3] = T_b[i * N + v];
            }
            for (int i = 0; i < 2 * ns; i++) {
                crss[i] = tcrss[i * N + v];
                crssb[i] = crss_b[i * N + v];
            }
            for (int c = 0; c < 6; c++) t6[c] = tsig[c * N + v];
            adj::sym_to_b(t6, sigB);
            for (int c = 0; c < 3; c++) dW[c] = tdW[c * N + v];
            double G = tG[v], Gb = G_b[v];
            double c5[ADJ_NS][5];
            for (int s = 0; s < ns; s++) adj::schmid(T, P.bn[p][s], P.bb[p][s], c5[s]);
            double gd[ADJ_NS], ed[6], dd[6][6];
            adj::rates(ns, c5, crss, P.nrs[p], P.gam0[p], sigB, gd, ed, dd);
            double gdb[ADJ_NS];
            for (int s = 0; s < ADJ_NS; s++) gdb[s] = 0.0;
            adj::HardParams H;
            adj::OrientParams O;
            load_phase(P, p, H, O);
            if (P.harden) {
                double pb[3] = {0.0, 0.0, 0.0};
                adj::harden_bar(H, gd, G, crssb, Gb, gdb, pb);
                par[p * EVPGPU_NPARAM + EVPGPU_P_TAU1] += pb[0];
                par[p * EVPGPU_NPARAM + EVPGPU_P_THETA0] += pb[1];
                par[p * EVPGPU_NPARAM + EVPGPU_P_THETA1] += pb[2];
            }
            double dWb[3] = {0.0, 0.0, 0.0};
            if (P.orient) adj::orient_bar(O, T, dW, gd, Tb, gdb, dWb);
            // Eps^k = Eps^(k-1) + dt edot: the eps_p adjoint seeds edot
            double edb[6], sb[6], cb[ADJ_NS][5], nb[ADJ_NS];
            for (int c = 0; c < 6; c++) {
                edb[c] = P.dt * ep_b[c * N + v];
                sb[c] = P.seed[c];
            }
            for (int s = 0; s < ns; s++) {
                nb[s] = 0.0;
                for (int j = 0; j < 5; j++) cb[s][j] = 0.0;
            }
            adj::rates_bar(ns, c5, crss, P.nrs[p], P.gam0[p], sigB, gdb, edb, sb, cb, crssb, nb);
            for (int s = 0; s < ns; s++) {
                adj::schmid_bar(T, P.bn[p][s], P.bb[p][s], cb[s], Tb);
                par[p * EVPGPU_NPARAM + EVPGPU_P_NRS] += nb[s];
            }
            // tangent C_t = (S + dt D)^-1
            double S[6][6], C[6][6];
            const double *sg = sgr + 36LL * orient_idx[v];
            for (int i = 0; i < 36; i++) S[i / 6][i % 6] = sg[i];
            if (!adj::tangent(S, dd, P.dt, C)) {
                atomicAdd(bad, 1);
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j < 6; j++) C[i][j] = 0.0;
            }
            int q = 0;
            for (int i = 0; i < 6; i++)
                for (int j = i; j < 6; j++) Ct[(q++) * N + v] = C[i][j];
            for (int i = 0; i < 6; i++) {
                double a = 0.0;
                for (int j = 0; j < 6; j++) a += C[i][j] * sb[j];
                sig_b[i * N + v] = sb[i];
                rhs[i * N + v] = a;
            }
            for (int c = 0; c < 3; c++) dW_b[c * N + v] = dWb[c];
            for (int i = 0; i < 9; i++) T_b[i * N + v] = Tb[i / 3][i % 3];
            for (int i = 0; i < 2 * ns; i++) crss_b[i * N + v] = crssb[i];
            G_b[v] = Gb;
        }
    }
    for (int i = 0; i < NPAR; i++) block_sum_write(par[i], partials, i, NPAR);
}

__device__ __forceinline__ void ct_apply(const double *__restrict__ Ct, long long N, long long v,
                                         const double x[6], double y[6])
{
    double C[6][6];
    int q = 0;
    for (int i = 0; i < 6; i++)
        for (int j = i; j < 6; j++) {
            double c = Ct[(q++) * N + v];
            C[i][j] = c;
            C[j][i] = c;
        }
    for (int i = 0; i < 6; i++) {
        double a = 0.0;
        for (int j = 0; j < 6; j++) a += C[i][j] * x[j];
        y[i] = a;
    }
}

// r_bar = C_t (lambda - sigma_bar): the adjoint of the solve's right-hand
// side; feeds Eps of the previous increment and, through edot, (T, crss, n).
__global__ __launch_bounds__(AB) void k_adj_post(AdjParams P, long long N, const int *__restrict__ phase,
                                                 const double *__restrict__ tape,
                                                 const double *__restrict__ Ct,
                                                 const double *__restrict__ sig_b,
                                                 const double *__restrict__ lam, double *__restrict__ ep_b,
                                                 double *__restrict__ T_b, double *__restrict__ crss_b,
                                                 double *__restrict__ partials)
{
    const long long v = (long long)blockIdx.x * AB + threadIdx.x;
    double par[NPAR];
    for (int i = 0; i < NPAR; i++) par[i] = 0.0;
    if (v < N) {
        int p = phase[v] - 1;
        p = p < 0 ? 0 : (p > 1 ? 1 : p);
        if (!P.is_gas[p]) {
            const int ns = P.nsys[p];
            const double *tsig = tape, *ttran = tape + 6 * N, *tcrss = tape + 15 * N;
            double d[6], rb[6];
            for (int c = 0; c < 6; c++) d[c] = lam[c * N + v] - sig_b[c * N + v];
            ct_apply(Ct, N, v, d, rb);
            double edb[6];
            for (int c = 0; c < 6; c++) {
                ep_b[c * N + v] += rb[c];
                edb[c] = P.dt * rb[c];
            }
            double T[3][3], Tb[3][3], crss[2 * ADJ_NS], crssb[2 * ADJ_NS], t6[6], sigB[6];
            for (int i = 0; i < 9; i++) {
                T[i / 3][i % 3] = ttran[i * N + v];
                Tb[i / 3][i % 3] = T_b[i * N + v];
            }
            for (int i = 0; i < 2 * ns; i++) {
                crss[i] = tcrss[i * N + v];
                crssb[i] = crss_b[i * N + v];
            }
            for (int c = 0; c < 6; c++) t6[c] = tsig[c * N + v];
            adj::sym_to_b(t6, sigB);
            double c5[ADJ_NS][5], cb[ADJ_NS][5], nb[ADJ_NS], sb[6] = {0, 0, 0, 0, 0, 0};
            for (int s = 0; s < ns; s++) {
                adj::schmid(T, P.bn[p][s], P.bb[p][s], c5[s]);
                nb[s] = 0.0;
                for (int j = 0; j < 5; j++) cb[s][j] = 0.0;
            }
            adj::rates_bar(ns, c5, crss, P.nrs[p], P.gam0[p], sigB, nullptr, edb, sb, cb, crssb, nb);
            for (int s = 0; s < ns; s++) {
                adj::schmid_bar(T, P.bn[p][s], P.bb[p][s], cb[s], Tb);
                par[p * EVPGPU_NPARAM + EVPGPU_P_NRS] += nb[s];
            }
            for (int i = 0; i < 9; i++) T_b[i * N + v] = Tb[i / 3][i % 3];
            for (int i = 0; i < 2 * ns; i++) crss_b[i * N + v] = crssb[i];
        }
    }
    for (int i = 0; i < NPAR; i++) block_sum_write(par[i], partials, i, NPAR);
}

// d L / d tau0: the initial CRSS adjoint, forward and backward, per phase
__global__ __launch_bounds__(AB) void k_adj_tau0(long long N, const int *__restrict__ phase, AdjParams P,
                                                 const double *__restrict__ crss_b,
                                                 double *__restrict__ partials)
{
    const long long v = (long long)blockIdx.x * AB + threadIdx.x;
    double par[NPAR];
    for (int i = 0; i < NPAR; i++) par[i] = 0.0;
    if (v < N) {
        int p = phase[v] - 1;
        p = p < 0 ? 0 : (p > 1 ? 1 : p);
        if (!P.is_gas[p]) {
            double s = 0.0;
            for (int i = 0; i < 2 * P.nsys[p]; i++) s += crss_b[i * N + v];
            par[p * EVPGPU_NPARAM + EVPGPU_P_TAU0] = s;
        }
    }
    for (int i = 0; i < NPAR; i++) block_sum_write(par[i], partials, i, NPAR);
}

// ---------------------------------------------------- spectral operators --
struct SpecParams {
    double c0[81];
    int free_c[6]; // 1 where the macroscopic strain component is unknown (stress-controlled)
};

// -Omega(n) sigma n : the vector u of the Green output -u x n
__device__ void omega_sig_n(const double *c0, const double n[3], const double s[3][3], double u[3])
{
    double A[3][3];
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) {
            double a = 0.0;
            for (int j = 0; j < 3; j++)
                for (int l = 0; l < 3; l++) a += c0[((i * 3 + j) * 3 + k) * 3 + l] * n[j] * n[l];
            A[i][k] = a;
        }
    double O[3][3];
    O[0][0] = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    O[0][1] = A[0][2] * A[2][1] - A[0][1] * A[2][2];
    O[0][2] = A[0][1] * A[1][2] - A[0][2] * A[1][1];
    O[1][0] = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    O[1][1] = A[0][0] * A[2][2] - A[0][2] * A[2][0];
    O[1][2] = A[0][2] * A[1][0] - A[0][0] * A[1][2];
    O[2][0] = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    O[2][1] = A[0][1] * A[2][0] - A[0][0] * A[2][1];
    O[2][2] = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    double det = A[0][0] * O[0][0] + A[0][1] * O[1][0] + A[0][2] * O[2][0];
    double t[3];
    for (int k = 0; k < 3; k++) t[k] = s[k][0] * n[0] + s[k][1] * n[1] + s[k][2] * n[2];
    for (int i = 0; i < 3; i++) u[i] = (O[i][0] * t[0] + O[i][1] * t[1] + O[i][2] * t[2]) / det;
}

// symmetric tensor (11 22 33 23 13 12) <-> full 3x3
__device__ __forceinline__ void t6_full(const double t[6], double m[3][3])
{
    m[0][0] = t[0]; m[1][1] = t[1]; m[2][2] = t[2];
    m[1][2] = m[2][1] = t[3];
    m[0][2] = m[2][0] = t[4];
    m[0][1] = m[1][0] = t[5];
}

// Projection onto admissible strains of a real-linear spectrum x (B-basis,
// re or im part), at a wave vector of type (id0, split, n1, n2)
__device__ void project_part(double xB[6], bool id0, bool split, const double n1[3], const double n2[3],
                             const int free_c[6])
{
    double e[6];
    adj::b_to_sym(xB, e);
    if (id0) {
        for (int c = 0; c < 6; c++)
            if (!free_c[c]) e[c] = 0.0;
    } else if (!split) {
        double E3[3][3], en[3];
        t6_full(e, E3);
        for (int i = 0; i < 3; i++) en[i] = E3[i][0] * n1[0] + E3[i][1] * n1[1] + E3[i][2] * n1[2];
        double nen = en[0] * n1[0] + en[1] * n1[1] + en[2] * n1[2];
        double a[3];
        for (int i = 0; i < 3; i++) a[i] = 2.0 * en[i] - nen * n1[i];
        const int ia[6] = {0, 1, 2, 1, 0, 0}, ib[6] = {0, 1, 2, 2, 2, 1};
        for (int c = 0; c < 6; c++) e[c] = 0.5 * (a[ia[c]] * n1[ib[c]] + a[ib[c]] * n1[ia[c]]);
    } else {
        // the admissible set is {sym(a x n1 + b x n2)}; its complement is m x m
        double m[3] = {n1[1] * n2[2] - n1[2] * n2[1], n1[2] * n2[0] - n1[0] * n2[2],
                       n1[0] * n2[1] - n1[1] * n2[0]};
        double mm = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        for (int i = 0; i < 3; i++) m[i] /= mm;
        double E3[3][3];
        t6_full(e, E3);
        double s = 0.0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) s += m[i] * E3[i][j] * m[j];
        e[0] -= s * m[0] * m[0];
        e[1] -= s * m[1] * m[1];
        e[2] -= s * m[2] * m[2];
        e[3] -= s * m[1] * m[2];
        e[4] -= s * m[0] * m[2];
        e[5] -= s * m[0] * m[1];
    }
    adj::sym_to_b(e, xB);
}

// Transpose of the map from the admissible strain change to the
// antisymmetric increment dW (23 13 12), applied to a real-linear part of
// the dW adjoint spectrum; returns the strain adjoint in the B-basis.
__device__ void skew_T_part(const double *c0, const double wb[3], bool id0, bool split, const double n1[3],
                            const double n2[3], const int free_c[6], double eB[6])
{
    double et[6] = {0, 0, 0, 0, 0, 0};
    if (id0) {
        // uniform part: dW_ij = dEbar_ij for the free shear components
        for (int c = 0; c < 3; c++)
            if (free_c[3 + c]) et[3 + c] = wb[c];
        adj::b_to_sym_T(et, eB);
        return;
    }
    if (!split) {
        // dW = skew(a x n), a = 2 e.n - (n.e.n) n  ->  a_bar = Om_bar n with
        // Om_bar_ij = wb/2 on (i,j), -wb/2 on (j,i); e_bar = 2 sym(a_bar x n)
        double Ob[3][3] = {{0, 0.5 * wb[2], 0.5 * wb[1]}, {-0.5 * wb[2], 0, 0.5 * wb[0]},
                           {-0.5 * wb[1], -0.5 * wb[0], 0}};
        double ab[3];
        for (int i = 0; i < 3; i++) ab[i] = Ob[i][0] * n1[0] + Ob[i][1] * n1[1] + Ob[i][2] * n1[2];
        // plain partials w.r.t. the component values (shears count both entries)
        et[0] = 2.0 * ab[0] * n1[0];
        et[1] = 2.0 * ab[1] * n1[1];
        et[2] = 2.0 * ab[2] * n1[2];
        et[3] = 2.0 * (ab[1] * n1[2] + ab[2] * n1[1]);
        et[4] = 2.0 * (ab[0] * n1[2] + ab[2] * n1[0]);
        et[5] = 2.0 * (ab[0] * n1[1] + ab[1] * n1[0]);
        adj::b_to_sym_T(et, eB);
        return;
    }
    // two directions: build G (sym part) and K (skew part) of
    // sigma_B -> -1/2 sum_i (Omega_i sigma n_i) x n_i, then e_bar = G^+ K^T wb
    double G[6][6], K[3][6];
    for (int j = 0; j < 6; j++) {
        double vb[6] = {0, 0, 0, 0, 0, 0}, st[6], S3[3][3], u1[3], u2[3], H[3][3];
        vb[j] = 1.0;
        adj::b_to_sym(vb, st);
        t6_full(st, S3);
        omega_sig_n(c0, n1, S3, u1);
        omega_sig_n(c0, n2, S3, u2);
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) H[a][b] = -0.5 * (u1[a] * n1[b] + u2[a] * n2[b]);
        double ht[6] = {H[0][0], H[1][1], H[2][2], 0.5 * (H[1][2] + H[2][1]), 0.5 * (H[0][2] + H[2][0]),
                        0.5 * (H[0][1] + H[1][0])};
        double hB[6];
        adj::sym_to_b(ht, hB);
        for (int i = 0; i < 6; i++) G[i][j] = hB[i];
        K[0][j] = 0.5 * (H[1][2] - H[2][1]);
        K[1][j] = 0.5 * (H[0][2] - H[2][0]);
        K[2][j] = 0.5 * (H[0][1] - H[1][0]);
    }
    double m[3] = {n1[1] * n2[2] - n1[2] * n2[1], n1[2] * n2[0] - n1[0] * n2[2], n1[0] * n2[1] - n1[1] * n2[0]};
    double mm = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    for (int i = 0; i < 3; i++) m[i] /= mm;
    double mt[6] = {m[0] * m[0], m[1] * m[1], m[2] * m[2], m[1] * m[2], m[0] * m[2], m[0] * m[1]}, vB[6];
    adj::sym_to_b(mt, vB);
    // M = v v^T - G is symmetric positive definite; solve M z = -(K^T wb)
    double M[6][6], y[6];
    for (int i = 0; i < 6; i++) {
        y[i] = -(K[0][i] * wb[0] + K[1][i] * wb[1] + K[2][i] * wb[2]);
        for (int j = 0; j < 6; j++) M[i][j] = vB[i] * vB[j] - 0.5 * (G[i][j] + G[j][i]);
    }
    for (int j = 0; j < 6; j++) {
        double d = M[j][j];
        for (int k = 0; k < j; k++) d -= M[j][k] * M[j][k];
        d = sqrt(d > 0.0 ? d : 1e-300);
        M[j][j] = d;
        for (int i = j + 1; i < 6; i++) {
            double s = M[i][j];
            for (int k = 0; k < j; k++) s -= M[i][k] * M[j][k];
            M[i][j] = s / d;
        }
    }
    for (int i = 0; i < 6; i++) {
        double s = y[i];
        for (int k = 0; k < i; k++) s -= M[i][k] * y[k];
        y[i] = s / M[i][i];
    }
    for (int i = 5; i >= 0; i--) {
        double s = y[i];
        for (int k = i + 1; k < 6; k++) s -= M[k][i] * y[k];
        y[i] = s / M[i][i];
    }
    for (int i = 0; i < 6; i++) eB[i] = y[i];
}

// spec6 <- P(spec6) [+ skew^T(w3)]  (B-basis, FP64 half spectrum)
__global__ __launch_bounds__(256) void k_adj_spectral(double2 *__restrict__ spec6, const double2 *__restrict__ w3,
                                                      SpecParams SP, int nx, int ny, int nz, int nzh,
                                                      long long NK)
{
    __shared__ double c0s[81];
    if (threadIdx.x < 81) c0s[threadIdx.x] = SP.c0[threadIdx.x];
    __syncthreads();
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= NK) return;
    const int kz = (int)(id % nzh);
    const int y = (int)((id / nzh) % ny);
    const int x = (int)(id / ((long long)nzh * ny));
    const bool id0 = (x == 0 && y == 0 && kz == 0);
    double n1[3] = {0, 0, 0}, n2[3] = {0, 0, 0};
    bool split = false;
    if (!id0) split = kspace::mode_dirs(x, y, kz, nx, ny, nz, n1, n2);
    double re[6], im[6];
    for (int c = 0; c < 6; c++) {
        double2 z = spec6[c * NK + id];
        re[c] = z.x;
        im[c] = z.y;
    }
    project_part(re, id0, split, n1, n2, SP.free_c);
    project_part(im, id0, split, n1, n2, SP.free_c);
    if (w3) {
        double wr[3], wi[3], er[6], ei[6];
        for (int c = 0; c < 3; c++) {
            double2 z = w3[c * NK + id];
            wr[c] = z.x;
            wi[c] = z.y;
        }
        skew_T_part(c0s, wr, id0, split, n1, n2, SP.free_c, er);
        skew_T_part(c0s, wi, id0, split, n1, n2, SP.free_c, ei);
        for (int c = 0; c < 6; c++) {
            re[c] += er[c];
            im[c] += ei[c];
        }
    }
    for (int c = 0; c < 6; c++) spec6[c * NK + id] = make_double2(re[c], im[c]);
}

// ------------------------------------------------------------------ CG --
// scal: 0 rr, 1 pq, 2 rr_new, 3 unused, 4 beta, 5 rr0
__global__ __launch_bounds__(AB) void k_cg_pct(long long N, const double *__restrict__ Ct,
                                               const double *__restrict__ r, double *__restrict__ p,
                                               double *__restrict__ t, const double *__restrict__ scal,
                                               int first, double *__restrict__ partials)
{
    const long long v = (long long)blockIdx.x * AB + threadIdx.x;
    double s = 0.0;
    if (v < N) {
        const double beta = first ? 0.0 : scal[4];
        double pv[6], tv[6];
        for (int c = 0; c < 6; c++) {
            double a = r[c * N + v] + (first ? 0.0 : beta * p[c * N + v]);
            pv[c] = a;
            p[c * N + v] = a;
        }
        ct_apply(Ct, N, v, pv, tv);
        for (int c = 0; c < 6; c++) {
            t[c * N + v] = tv[c];
            s += pv[c] * tv[c];
        }
    }
    block_sum_write(s, partials, 0, 1);
}

__global__ __launch_bounds__(AB) void k_cg_xr(long long N, const double *__restrict__ p,
                                              const double *__restrict__ q, double *__restrict__ x,
                                              double *__restrict__ r, const double *__restrict__ scal,
                                              double *__restrict__ partials)
{
    const long long v = (long long)blockIdx.x * AB + threadIdx.x;
    double s = 0.0;
    if (v < N) {
        const double alpha = scal[1] != 0.0 ? scal[0] / scal[1] : 0.0;
        for (int c = 0; c < 6; c++) {
            x[c * N + v] += alpha * p[c * N + v];
            double rv = r[c * N + v] - alpha * q[c * N + v];
            r[c * N + v] = rv;
            s += rv * rv;
        }
    }
    block_sum_write(s, partials, 0, 1);
}

__global__ __launch_bounds__(AB) void k_dot6(long long N, const double *__restrict__ a,
                                             const double *__restrict__ b, double *__restrict__ partials)
{
    const long long v = (long long)blockIdx.x * AB + threadIdx.x;
    double s = 0.0;
    if (v < N)
        for (int c = 0; c < 6; c++) s += a[c * N + v] * b[c * N + v];
    block_sum_write(s, partials, 0, 1);
}

// mode 0: rr = rr0 = rr_new ; mode 1: beta = rr_new / rr, rr = rr_new
__global__ void k_cg_scal(double *scal, int mode)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double rn = scal[2];
    if (mode == 0) {
        scal[0] = rn;
        scal[5] = rn;
        scal[4] = 0.0;
    } else {
        scal[4] = scal[0] != 0.0 ? rn / scal[0] : 0.0;
        scal[0] = rn;
    }
}

__global__ void k_sub3(long long n, const double *__restrict__ a, const double *__restrict__ b,
                       double *__restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] - b[i];
}

// ----------------------------------------------------------- host side --
#define ACK(x)                                    \
    do {                                          \
        int s_ = (x);                             \
        if (s_) return s_;                        \
    } while (0)

template <typename T> int dalloc(T **p, size_t n)
{
    if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return EVPGPU_ERR_NOMEM;
    }
    return EVPGPU_OK;
}

unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

int launch_ok() { return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP; }

// the state fields a checkpoint holds, in order, with their component counts
struct Field {
    double *d;
    int nc;
};

void state_fields(evpgpu_ctx *c, Field f[9])
{
    Field g[9] = {{c->sig, 6}, {c->E, 6}, {c->W, 3}, {c->epsp, 6}, {c->edot, 6},
                  {c->tran, 9}, {c->crss, 24}, {c->gamacum, 1}, {c->gamdot, 12}};
    for (int i = 0; i < 9; i++) f[i] = g[i];
}

int snapshot(evpgpu_ctx *c, int slot)
{
    evpgpu_adj *A = c->adj;
    const long long N = c->N;
    Field f[9];
    state_fields(c, f);
    double *dst = A->cp + (long long)slot * STATE_DBL * N;
    for (int i = 0; i < 9; i++) {
        EVP_CHECK(hipMemcpyAsync(dst, f[i].d, sizeof(double) * f[i].nc * N, hipMemcpyDeviceToDevice, c->stream));
        dst += (long long)f[i].nc * N;
    }
    A->cp_m[slot] = c->m;
    memcpy(A->cp_dga + 9 * slot, c->disgrad_avg_actual, sizeof(double) * 9);
    return EVPGPU_OK;
}

int restore(evpgpu_ctx *c, int slot)
{
    evpgpu_adj *A = c->adj;
    const long long N = c->N;
    Field f[9];
    state_fields(c, f);
    const double *src = A->cp + (long long)slot * STATE_DBL * N;
    for (int i = 0; i < 9; i++) {
        EVP_CHECK(hipMemcpyAsync(f[i].d, src, sizeof(double) * f[i].nc * N, hipMemcpyDeviceToDevice, c->stream));
        src += (long long)f[i].nc * N;
    }
    c->m = A->cp_m[slot];
    memcpy(c->disgrad_avg_actual, A->cp_dga + 9 * slot, sizeof(double) * 9);
    return EVPGPU_OK;
}

AdjParams adj_params(const evpgpu_ctx *c, int k, const double *seed_t)
{
    const evpgpu_adj *A = c->adj;
    AdjParams P;
    memset(&P, 0, sizeof(P));
    P.dt = A->ld.dt;
    memcpy(P.udot_a, c->udot_a, sizeof(P.udot_a));
    P.orient = A->f_orient[k];
    P.harden = A->f_harden[k];
    for (int p = 0; p < 2; p++) {
        if (p >= A->nph) {
            P.is_gas[p] = 1;
            continue;
        }
        const evpgpu_phase &ph = A->ph[p];
        P.is_gas[p] = ph.is_gas;
        P.nsys[p] = ph.nsys;
        for (int s = 0; s < ADJ_NS; s++) {
            P.nrs[p][s] = ph.nrs[s];
            P.gam0[p][s] = ph.gam0[s];
            for (int j = 0; j < ADJ_NS; j++) P.hard[p][s][j] = ph.hard[s][j];
            for (int j = 0; j < 3; j++) {
                P.tau[p][s][j] = ph.tau[s][j];
                P.bn[p][s][j] = ph.bn[s][j];
                P.bb[p][s][j] = ph.bb[s][j];
            }
            P.thet[p][s][0] = ph.thet[s][0];
            P.thet[p][s][1] = ph.thet[s][1];
        }
    }
    // Sigbar_c = (1/N) sum_v b_to_sym(sigma_B(v))_c  ->  seed_B = b_to_sym^T(ybar) / N
    double yb[6] = {0, 0, 0, 0, 0, 0};
    if (seed_t)
        for (int i = 0; i < 6; i++) yb[i] = seed_t[i] / (double)c->N;
    adj::b_to_sym_T(yb, P.seed);
    return P;
}

SpecParams spec_params(const evpgpu_ctx *c)
{
    SpecParams SP;
    memcpy(SP.c0, c->adj->ld.c0, sizeof(SP.c0));
    for (int i = 0; i < 6; i++) SP.free_c[i] = c->adj->ld.stress_bc_flag[i] != 0;
    return SP;
}

// sum of the NPAR block partials into g (host)
int add_partials(evpgpu_ctx *c, int nb, double *g)
{
    ACK(reduce_partials(c, c->partials, nb, NPAR, c->red));
    EVP_CHECK(hipMemcpyAsync(c->h_red, c->red, NPAR * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    EVP_CHECK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < NPAR; i++) g[i] += c->h_red[i];
    return EVPGPU_OK;
}

int spectral(evpgpu_ctx *c, const double2 *w3)
{
    SpecParams SP = spec_params(c);
    hipLaunchKernelGGL(k_adj_spectral, dim3(nblk(c->NK, 256)), dim3(256), 0, c->stream, (double2 *)c->spec, w3,
                       SP, c->nx, c->ny, c->nz, c->nzh, c->NK);
    return launch_ok();
}

// A x = P (C_t x) applied to p (already admissible), result in q
int cg_solve(evpgpu_ctx *c, double tol, int cg_max, int *iters, double *rel)
{
    evpgpu_adj *A = c->adj;
    const long long N = c->N;
    const unsigned nb = nblk(N, AB);
    const double scale = 1.0 / (double)N;
    *iters = 0;
    *rel = 0.0;
    ACK(fill_zero(c, A->x, 6 * N));
    hipLaunchKernelGGL(k_dot6, dim3(nb), dim3(AB), 0, c->stream, N, A->r, A->r, c->partials);
    ACK(launch_ok());
    ACK(reduce_partials(c, c->partials, nb, 1, A->scal + 2));
    hipLaunchKernelGGL(k_cg_scal, dim3(1), dim3(64), 0, c->stream, A->scal, 0);
    ACK(launch_ok());
    EVP_CHECK(hipMemcpyAsync(A->h_scal, A->scal, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    EVP_CHECK(hipStreamSynchronize(c->stream));
    const double rr0 = A->h_scal[5];
    if (!(rr0 > 0.0)) return EVPGPU_OK;
    for (int it = 0; it < cg_max; it++) {
        hipLaunchKernelGGL(k_cg_pct, dim3(nb), dim3(AB), 0, c->stream, N, A->Ct, A->r, A->p, A->t, A->scal,
                           it == 0, c->partials);
        ACK(launch_ok());
        ACK(reduce_partials(c, c->partials, nb, 1, A->scal + 1));
        ACK(fft_forward(c, A->t, 6, EVPGPU_FP64));
        ACK(spectral(c, nullptr));
        ACK(fft_inverse_real(c, (double2 *)c->spec, 6, A->q, scale));
        hipLaunchKernelGGL(k_cg_xr, dim3(nb), dim3(AB), 0, c->stream, N, A->p, A->q, A->x, A->r, A->scal,
                           c->partials);
        ACK(launch_ok());
        ACK(reduce_partials(c, c->partials, nb, 1, A->scal + 2));
        hipLaunchKernelGGL(k_cg_scal, dim3(1), dim3(64), 0, c->stream, A->scal, 1);
        ACK(launch_ok());
        *iters = it + 1;
        if ((it + 1) % 8 == 0 || it + 1 == cg_max) {
            EVP_CHECK(hipMemcpyAsync(A->h_scal, A->scal, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            EVP_CHECK(hipStreamSynchronize(c->stream));
            *rel = sqrt(A->h_scal[0] / rr0);
            if (*rel <= tol) break;
        }
    }
    return EVPGPU_OK;
}

// one increment of the reverse pass; tape slot holds its inputs
int adjoint_step(evpgpu_ctx *c, int k, const double *tape, const double *seed_t, double tol, int cg_max,
                 double *g, evpgpu_adjoint_stats *S)
{
    evpgpu_adj *A = c->adj;
    const long long N = c->N;
    const unsigned nb = nblk(N, AB);
    AdjParams P = adj_params(c, k, seed_t);
    hipLaunchKernelGGL(k_adj_pre, dim3(nb), dim3(AB), 0, c->stream, P, N, c->phase, c->orient, c->sgr, tape,
                       A->ep_b, A->T_b, A->crss_b, A->G_b, A->sig_b, A->Ct, A->t, A->dWb, c->partials,
                       A->d_bad);
    ACK(launch_ok());
    ACK(add_partials(c, nb, g));
    // right-hand side g = P (C_t sigma_bar) + skew^T(dW_bar_k - dW_bar_k+1)
    ACK(fft_forward(c, A->t, 6, EVPGPU_FP64));
    const bool wpath = A->f_orient[k] || (k + 1 < A->n && A->f_orient[k + 1]);
    if (wpath) {
        hipLaunchKernelGGL(k_sub3, dim3(nblk(3 * N, 256)), dim3(256), 0, c->stream, 3 * N, A->dWb, A->dWb_next,
                           c->rwork);
        ACK(launch_ok());
        ACK(fft_forward_to(c, c->rwork, 3, c->acc));
    }
    ACK(spectral(c, wpath ? c->acc : nullptr));
    ACK(fft_inverse_real(c, (double2 *)c->spec, 6, A->r, 1.0 / (double)N));
    int it = 0;
    double rel = 0.0;
    ACK(cg_solve(c, tol, cg_max, &it, &rel));
    S->cg_iters += it;
    if (it > S->cg_max_iters) S->cg_max_iters = it;
    if (rel > S->cg_rel_resid) S->cg_rel_resid = rel;
    hipLaunchKernelGGL(k_adj_post, dim3(nb), dim3(AB), 0, c->stream, P, N, c->phase, tape, A->Ct, A->sig_b,
                       A->x, A->ep_b, A->T_b, A->crss_b, c->partials);
    ACK(launch_ok());
    ACK(add_partials(c, nb, g));
    double *tmp = A->dWb_next;
    A->dWb_next = A->dWb;
    A->dWb = tmp;
    return EVPGPU_OK;
}

} // namespace

void adjoint_free(evpgpu_ctx *c)
{
    if (!c || !c->adj) return;
    evpgpu_adj *A = c->adj;
    void *dev[] = {A->cp, A->tape, A->ep_b, A->T_b, A->crss_b, A->G_b, A->dWb, A->dWb_next, A->sig_b,
                   A->Ct, A->x, A->r, A->p, A->t, A->q, A->scal, A->d_bad};
    for (size_t i = 0; i < sizeof(dev) / sizeof(dev[0]); i++)
        if (dev[i]) (void)hipFree(dev[i]);
    if (A->h_scal) (void)hipHostFree(A->h_scal);
    free(A->iters);
    free(A->fp64);
    free(A->f_orient);
    free(A->f_harden);
    free(A->cp_m);
    free(A->cp_dga);
    free(A);
    c->adj = nullptr;
}

extern "C" {

int evpgpu_adjoint_record(evpgpu_ctx *c, int n, int ck, double *sig_avg, evpgpu_stats *stats)
{
    if (!c || n < 1 || ck < 1) return EVPGPU_ERR_ARG;
    if (!(c->have_phases && c->have_loading && c->have_ms && c->have_state)) return EVPGPU_ERR_STATE;
    for (int p = 0; p < c->nphases; p++)
        if (c->ph[p].is_gas) return EVPGPU_ERR_ARG;
    if (c->ld.elast_bc != 1 || c->creep_flag == 1) return EVPGPU_ERR_ARG;
    if (c->dd_on) return EVPGPU_ERR_STATE; // the reverse pass is of the power law
    adjoint_free(c);
    evpgpu_adj *A = (evpgpu_adj *)calloc(1, sizeof(evpgpu_adj));
    c->adj = A;
    A->n = n;
    A->ck = ck < n ? ck : n;
    A->ncp = (n + A->ck - 1) / A->ck;
    A->nph = c->nphases;
    memcpy(A->ph, c->ph, sizeof(A->ph));
    A->ld = c->ld;
    A->iters = (int *)calloc(n, sizeof(int));
    A->fp64 = (int *)calloc(n, sizeof(int));
    A->f_orient = (int *)calloc(n, sizeof(int));
    A->f_harden = (int *)calloc(n, sizeof(int));
    A->cp_m = (evpgpu_macro *)calloc(A->ncp + 1, sizeof(evpgpu_macro));
    A->cp_dga = (double *)calloc(9 * (A->ncp + 1), sizeof(double));
    const long long N = c->N;
    int st = dalloc(&A->cp, (size_t)(A->ncp + 1) * STATE_DBL * N);
    if (!st) st = dalloc(&A->tape, (size_t)A->ck * TAPE_DBL * N);
    double **f6[] = {&A->ep_b, &A->sig_b, &A->x, &A->r, &A->p, &A->t, &A->q};
    for (int i = 0; i < 7 && !st; i++) st = dalloc(f6[i], 6 * N);
    if (!st) st = dalloc(&A->T_b, 9 * N);
    if (!st) st = dalloc(&A->crss_b, 24 * N);
    if (!st) st = dalloc(&A->G_b, N);
    if (!st) st = dalloc(&A->dWb, 3 * N);
    if (!st) st = dalloc(&A->dWb_next, 3 * N);
    if (!st) st = dalloc(&A->Ct, 21 * N);
    if (!st) st = dalloc(&A->scal, 8);
    if (!st) st = dalloc(&A->d_bad, 1);
    if (!st && hipHostMalloc((void **)&A->h_scal, 8 * sizeof(double), 0) != hipSuccess) st = EVPGPU_ERR_NOMEM;
    if (st) {
        adjoint_free(c);
        return st;
    }
    for (int k = 0; k < n; k++) {
        if (k % A->ck == 0) ACK(snapshot(c, k / A->ck));
        A->f_orient[k] = (c->ld.update_flag == 1 && c->m.istep > 1);
        A->f_harden[k] = (c->ld.hard_flag == 1 && c->m.istep > 1);
        evpgpu_stats S;
        ACK(increment_core(c, 0, 0, -1, &S));
        if (S.stopped) return EVPGPU_ERR_STATE;
        A->iters[k] = S.iters;
        A->fp64[k] = S.fp64_iters;
        if (stats) stats[k] = S;
        if (sig_avg) {
            const double *s = c->m.sig_avg;
            double *o = sig_avg + 6 * k;
            o[0] = s[0]; o[1] = s[4]; o[2] = s[8];
            o[3] = s[5]; o[4] = s[2]; o[5] = s[1];
        }
    }
    ACK(snapshot(c, A->ncp));
    EVP_CHECK(hipStreamSynchronize(c->stream));
    return EVPGPU_OK;
}

int evpgpu_adjoint_run(evpgpu_ctx *c, const double *sig_avg_bar, double tol, int cg_max, double *grad,
                       evpgpu_adjoint_stats *stats)
{
    if (!c || !c->adj || !sig_avg_bar || !grad || cg_max < 1 || !(tol > 0.0)) return EVPGPU_ERR_ARG;
    evpgpu_adj *A = c->adj;
    const long long N = c->N;
    const double t0 = now_s();
    double replay = 0.0;
    evpgpu_adjoint_stats S;
    memset(&S, 0, sizeof(S));
    double g[NPAR];
    for (int i = 0; i < NPAR; i++) g[i] = 0.0;
    double *zero[] = {A->ep_b, A->T_b, A->crss_b, A->G_b, A->dWb, A->dWb_next};
    const int nz[] = {6, 9, 24, 1, 3, 3};
    for (int i = 0; i < 6; i++) ACK(fill_zero(c, zero[i], nz[i] * N));
    EVP_CHECK(hipMemsetAsync(A->d_bad, 0, sizeof(int), c->stream));
    for (int seg = A->ncp - 1; seg >= 0; seg--) {
        const int k0 = seg * A->ck, k1 = (k0 + A->ck < A->n) ? k0 + A->ck : A->n;
        const double tr = now_s();
        ACK(restore(c, seg));
        for (int k = k0; k < k1; k++) {
            double *slot = A->tape + (long long)(k - k0) * TAPE_DBL * N;
            EVP_CHECK(hipMemcpyAsync(slot + 6 * N, c->tran, 9 * N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            EVP_CHECK(hipMemcpyAsync(slot + 15 * N, c->crss, 24 * N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            EVP_CHECK(hipMemcpyAsync(slot + 39 * N, c->gamacum, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            evpgpu_stats R;
            ACK(increment_core(c, 0, A->iters[k] - A->fp64[k], A->fp64[k], &R));
            EVP_CHECK(hipMemcpyAsync(slot, c->sig, 6 * N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            EVP_CHECK(hipMemcpyAsync(slot + 40 * N, c->rwork, 3 * N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        }
        EVP_CHECK(hipStreamSynchronize(c->stream));
        replay += now_s() - tr;
        for (int k = k1 - 1; k >= k0; k--) {
            const double *slot = A->tape + (long long)(k - k0) * TAPE_DBL * N;
            ACK(adjoint_step(c, k, slot, sig_avg_bar + 6 * k, tol, cg_max, g, &S));
        }
    }
    {
        AdjParams P = adj_params(c, 0, nullptr);
        const unsigned nb = nblk(N, AB);
        hipLaunchKernelGGL(k_adj_tau0, dim3(nb), dim3(AB), 0, c->stream, N, c->phase, P, A->crss_b, c->partials);
        ACK(launch_ok());
        ACK(add_partials(c, nb, g));
    }
    int bad = 0;
    EVP_CHECK(hipMemcpy(&bad, A->d_bad, sizeof(int), hipMemcpyDeviceToHost));
    ACK(restore(c, A->ncp));
    EVP_CHECK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < NPAR; i++) grad[i] = g[i];
    S.seconds = now_s() - t0;
    S.replay_seconds = replay;
    if (stats) *stats = S;
    return bad ? EVPGPU_ERR_SINGULAR : EVPGPU_OK;
}

} // extern "C"
