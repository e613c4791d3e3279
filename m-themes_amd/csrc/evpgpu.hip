// evpgpu.hip -- C ABI of the MI355X EVPFFT loop: context, state transfer and
// the loading increment (src/evolution.c:1318-1547 + reconstructed harden).
//
// Per Green iteration (src/evolution.c:1355-1462) the device runs
//   R2C(sigma, 6 comps) -> Gamma_eff (in place, antisym part accumulated)
//   -> C2R fused with  E += sym(dDisGradAvg) + re/N
//   -> per-voxel Newton update + block partials -> one deterministic row sum
// -> get_smacro (src/evolution.c:473-547), the error normalisation and the
// loop test in a one-workgroup kernel (loopctl.hip).  The control stays on
// the device, so the host queues iterations in batches and reads the loop's
// flags once per batch, while the device runs the next one.
//
// Mixed precision splits sigma = sigma_A + delta.  Every anchor_every
// iterations sigma_A = fp32(sigma) is transformed in FP64 and its Green
// correction kept in FP32 (real space, cA, and the antisymmetric spectrum,
// skA); in between only delta, which shrinks as the loop converges, passes
// through the FP32 FFT/Green/C2R path, so the FP32 rounding scales with the
// per-iteration update rather than with |sigma|.  Newton, E, the reductions
// and the antisymmetric accumulator stay FP64.  When the mixed loop has
// converged, the strain change of the increment is projected in FP64 onto
// compatible fields (removing the rounding that Gamma can never correct),
// then FP64 sweeps run until the FP64 residual converges again (at least
// one sweep).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "evpgpu_impl.h"

// ------------------------------------------------------------ host tensors --
// ---- kernel timing ----
void tmark(evpgpu_ctx *c, int k, int end)
{
    if (!c->timing || c->tskip || !((c->timing_groups >> k) & 1u)) return;
    (void)hipEventRecord(c->ev[c->tslot][k][end], c->stream);
    if (end) c->tmask[c->tslot] |= 1u << k;
}

// add the elapsed times of the event pairs recorded in a slot since its last
// collect (once the device has passed them), or drop them (keep == 0: the
// launches of an iteration queued past the end of the loop, which returned
// at once); a pair never recorded is never queried, so no error is left
// behind for the next launch check
void tcollect_slot(evpgpu_ctx *c, int slot, int keep)
{
    if (!c->timing) return;
    for (int k = 0; k < EVP_NTIMERS && keep; k++) {
        if (!(c->tmask[slot] & (1u << k))) continue;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev[slot][k][0], c->ev[slot][k][1]) == hipSuccess) {
            c->t_ms[k] += ms;
            c->t_n[k]++;
        } else {
            (void)hipGetLastError();
        }
    }
    c->tmask[slot] = 0;
}

void tkernel(evpgpu_ctx *c, int k, hipEvent_t *start, hipEvent_t *stop)
{
    *start = *stop = nullptr;
    if (!c->timing || c->tskip || !((c->timing_groups >> k) & 1u)) return;
    *start = c->ev[c->tslot][k][0];
    *stop = c->ev[c->tslot][k][1];
    c->tmask[c->tslot] |= 1u << k;
}

// the current slot (after the stream has been synchronised)
void tcollect(evpgpu_ctx *c)
{
    tcollect_slot(c, c->tslot, 1);
}

namespace {

// VonMises (src/kinematics.c:465-484)
double vonmises(const double t[9])
{
    double tr = t[0] + t[4] + t[8], vm = 0.0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double d = (t[a * 3 + b] + t[b * 3 + a]) / 2. - (double)(a == b) * tr / 3.0;
            vm += d * d;
        }
    return sqrt(2. / 3. * vm);
}

void sym6_to_full(const double s[6], double t[9])
{
    t[0] = s[0]; t[4] = s[1]; t[8] = s[2];
    t[5] = t[7] = s[3];
    t[2] = t[6] = s[4];
    t[1] = t[3] = s[5];
}

double now_s()
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

// grid lengths: any length in [2, 512]; powers of two from 8 take the
// Stockham kernels, the others the direct-DFT ones (fft.hip)
bool len_ok(int n) { return n >= 2 && n <= 512; }

template <typename T> int dalloc(T **p, size_t n)
{
    if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return EVPGPU_ERR_NOMEM;
    }
    return EVPGPU_OK;
}

// smallest eigenvalue of a symmetric 6x6 matrix (cyclic Jacobi rotations)
double sym6_min_eig(double a[6][6])
{
    for (int sweep = 0; sweep < 50; sweep++) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < 6; i++) {
            diag += a[i][i] * a[i][i];
            for (int j = i + 1; j < 6; j++) off += a[i][j] * a[i][j];
        }
        if (off <= 1e-34 * diag) break;
        for (int p = 0; p < 6; p++)
            for (int q = p + 1; q < 6; q++) {
                if (a[p][q] == 0.0) continue;
                double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < 6; k++) {
                    double akp = a[k][p], akq = a[k][q];
                    a[k][p] = cs * akp - sn * akq;
                    a[k][q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < 6; k++) {
                    double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = cs * apk - sn * aqk;
                    a[q][k] = sn * apk + cs * aqk;
                }
            }
    }
    double m = a[0][0];
    for (int i = 1; i < 6; i++) m = a[i][i] < m ? a[i][i] : m;
    return m;
}

// min over orientations of lambda_min(S0 + S), both from their upper
// triangles as the Newton kernel reads them (bounds its step, see newton.hip)
void update_lam_min(evpgpu_ctx *c)
{
    c->lam_min = -1.0;
    if (!c->have_loading || !c->h_sgr) return;
    double m = 1e300;
    for (int o = 0; o < c->n_orient; o++) {
        double a[6][6];
        for (int i = 0; i < 6; i++)
            for (int j = i; j < 6; j++) {
                a[i][j] = c->ld.s066[i * 6 + j] + c->h_sgr[36LL * o + i * 6 + j];
                a[j][i] = a[i][j];
            }
        double e = sym6_min_eig(a);
        m = e < m ? e : m;
    }
    c->lam_min = m;
}

NewtonParams newton_params(const evpgpu_ctx *c)
{
    NewtonParams P;
    memset(&P, 0, sizeof(P));
    // C0 and S0 are symmetric: their upper triangles (S from the
    // per-orientation table is taken the same way)
    for (int i = 0; i < 6; i++)
        for (int j = i; j < 6; j++) {
            P.c0p[PK(i, j)] = c->ld.c066[i * 6 + j];
            P.s0p[PK(i, j)] = c->ld.s066[i * 6 + j];
        }
    P.dt = c->ld.dt;
    for (int p = 0; p < 2; p++) {
        if (p >= c->nphases) {
            P.is_gas[p] = 1;
            continue;
        }
        const evpgpu_phase &ph = c->ph[p];
        P.is_gas[p] = ph.is_gas;
        P.nsys[p] = ph.nsys;
        int integral = ph.nsys > 0;
        for (int s = 0; s < ph.nsys; s++) {
            P.nrs[p][s] = ph.nrs[s];
            P.gam0[p][s] = ph.gam0[s];
            double n = ph.nrs[s];
            if (!(n == floor(n) && n >= 1 && n <= 64 && n == ph.nrs[0])) integral = 0;
        }
        P.int_n[p] = integral ? (int)ph.nrs[0] : 0;
        memcpy(P.bn[p], ph.bn, sizeof(P.bn[p]));
        memcpy(P.bb[p], ph.bb, sizeof(P.bb[p]));
    }
    P.crss_sym = c->crss_sym;
    P.nr_tol = c->nr_tol;
    P.tolx = c->tolx;
    const double kt = c->lam_min > 0.0 ? 0.99 * c->tolx * c->lam_min : 0.0;
    P.kmin_tolx2 = kt * kt;
    P.dd = c->dd_on;
    P.dd_first = 0;
    return P;
}

// the loading constants of the loop control's get_smacro
CtlConst ctl_const(const evpgpu_ctx *c)
{
    CtlConst K;
    memset(&K, 0, sizeof(K));
    memcpy(K.s0, c->ld.s0, sizeof(K.s0));
    memcpy(K.scauchy, c->scauchy, sizeof(K.scauchy));
    K.wgt_ph1 = c->ld.wgt_ph1;
    K.scale = 1.0 / (double)c->N;
    memcpy(K.velgrad_bc, c->ld.velgrad_bc_flag, sizeof(K.velgrad_bc));
    memcpy(K.stress_bc, c->ld.stress_bc_flag, sizeof(K.stress_bc));
    return K;
}

int green_iteration_queue(evpgpu_ctx *c, int prec, int refresh, const CtlConst &K);

// One Green iteration, queued: prec FP64 runs the FFT/Green/C2R path in FP64
// on sigma; MIXED runs it in FP32 on delta = sigma - sigma_A (anchor
// refreshed first when refresh != 0).  The strain update adds the dD of the
// last control step: read from the control block by the fused kernels, from
// the host copy of it otherwise (the loop then waits for every iteration).
// Ends with the loop control's step.
int green_iteration(evpgpu_ctx *c, int prec, int refresh, const CtlConst &K)
{
    // the timers record one Green iteration in timing_every
    c->tskip = c->timing && (c->tcount++ % c->timing_every) != 0;
    int st = green_iteration_queue(c, prec, refresh, K);
    c->tskip = 0;
    return st;
}

int green_iteration_queue(evpgpu_ctx *c, int prec, int refresh, const CtlConst &K)
{
    const double scale = K.scale;
    int st = EVPGPU_OK;
    const bool fused = c->fused && fused_ok(c);
    if (prec == EVPGPU_MIXED && refresh) {
        tmark(c, 5, 0);
        st = to_f32_launch(c, c->sig, c->sigA, 6 * c->N);
        if (fused) {
            if (!st) st = fused_anchor(c, c->sigA, c->cA, scale);
        } else {
            if (!st) st = fft_forward_anchor(c, c->sigA);
            if (!st) st = green_anchor(c);
            if (!st) st = fft_inverse_anchor(c, c->cA, scale);
        }
        if (st) return st;
        tmark(c, 5, 1);
    }
    if (fused) {
        const double *addd = (const double *)((const char *)c->ctl + offsetof(LoopCtl, add));
        st = prec == EVPGPU_MIXED ? fused_iteration_delta(c, c->sig, c->sigA, c->E, addd, c->cA, scale)
                                  : fused_iteration_fp64(c, c->sig, c->E, addd, scale);
        if (st) return st;
    } else {
    const double *add = c->h_ctl[0].add;
    tmark(c, 1, 0);
    st = prec == EVPGPU_MIXED ? fft_forward_delta(c, c->sig, c->sigA) : fft_forward(c, c->sig, 6, prec);
    if (st) return st;
    tmark(c, 1, 1);
    tmark(c, 2, 0);
    st = prec == EVPGPU_MIXED ? green_delta(c) : green_apply(c, prec, 1);
    if (st) return st;
    tmark(c, 2, 1);
    tmark(c, 3, 0);
    st = prec == EVPGPU_MIXED ? fft_inverse_delta(c, c->E, add, c->cA, scale)
                              : fft_inverse_update(c, 6, prec, c->E, add, scale);
    if (st) return st;
    tmark(c, 3, 1);
    }
    NewtonParams P = newton_params(c);
    P.dd_first = c->dd_first;
    P.skip = c->skip;
    c->dd_first = 0;
    int nblk = 0;
    hipEvent_t e0, e1;
    tkernel(c, 0, &e0, &e1);
    st = newton_launch(c, P, c->sig, c->sig_alt, c->E, c->epsp, c->edot, c->tran, c->dd_on ? c->dd_k : c->crss,
                       c->gamdot, c->phase, c->orient, c->sgr, c->N, c->partials, &nblk, c->d_status, nullptr,
                       nullptr, c->newton_shortcut, e0, e1);
    if (st) return st;
    {
        double *t = c->sig;
        c->sig = c->sig_alt;
        c->sig_alt = t;
    }
    return ctl_iter_launch(c, nblk);
}

// the phase-end test of the loop control (k_ctl_phase) on the host copy
bool phase_ends(const LoopCtl &L, int limit, int conv)
{
    return !(L.f.iter < limit && (!conv || fmax(fabs(L.err_s), fabs(L.err_e)) > L.tol));
}

// One phase of the Green loop: iterations of one precision while
// iter < limit (and, with conv, while the error norms exceed Err); first
// runs at least one.  With the fused kernels and loop_batch > 0 the
// iterations are queued in batches of loop_batch, two batches in flight: the
// host reads the older batch's flags while the device runs the newer one,
// and the iterations queued past the end of the phase return at once.
// Otherwise the host waits for every iteration.  On return the host copy of
// the control block (h_ctl[0]) is current and c->sig is the last stress.
int run_phase(evpgpu_ctx *c, int prec, int limit, int conv, int first, const CtlConst &K, int *imix)
{
    int st = ctl_phase_launch(c, limit, conv, first);
    if (st) return st;
    if (!first && phase_ends(c->h_ctl[0], limit, conv)) return EVPGPU_OK;
    const bool ahead = c->fused && fused_ok(c) && c->loop_batch > 0;
    const int B = ahead ? c->loop_batch : 1, depth = ahead ? 2 : 1;
    const int ae = c->anchor_every > 0 ? c->anchor_every : 1;
    const int iter0 = c->h_ctl[0].f.iter;
    c->skip = ahead ? (const int *)((const char *)c->ctl + offsetof(LoopCtl, f.done)) : nullptr;
    long long sent = 0, seen = 0, queued = 0;
    bool done = false;
    while (!done && !st) {
        while (!st && sent - seen < depth) {
            for (int b = 0; b < B && !st; b++) {
                c->tslot = (int)((sent * B + b) % EVP_NSLOT);
                const int refresh = prec == EVPGPU_MIXED && (*imix)++ % ae == 0;
                st = green_iteration(c, prec, refresh, K);
                if (!st) queued++;
            }
            if (st) break;
            LoopCtl *h = ahead ? &c->h_ctl[1 + sent % 2] : &c->h_ctl[0];
            const size_t bytes = ahead ? sizeof(LoopFlags) : sizeof(LoopCtl);
            if (hipMemcpyAsync(h, c->ctl, bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipEventRecord(c->bev[sent % 2], c->stream) != hipSuccess)
                st = EVPGPU_ERR_HIP;
            sent++;
        }
        if (st) break;
        if (hipEventSynchronize(c->bev[seen % 2]) != hipSuccess) {
            st = EVPGPU_ERR_HIP;
            break;
        }
        const LoopFlags f = (ahead ? c->h_ctl[1 + seen % 2] : c->h_ctl[0]).f;
        for (int b = 0; b < B; b++) {
            const long long j = seen * B + b;
            tcollect_slot(c, (int)(j % EVP_NSLOT), iter0 + j < f.iter);
        }
        seen++;
        done = f.done != 0;
    }
    // the batches still queued run nothing; wait for them, drop their timer
    // pairs (every slot not collected above, a batch cut short by an error
    // included) and bring the control block back
    if (hipStreamSynchronize(c->stream) != hipSuccess && !st) st = EVPGPU_ERR_HIP;
    for (int q = 0; q < EVP_NSLOT; q++) tcollect_slot(c, q, 0);
    c->tslot = 0;
    c->skip = nullptr;
    if (st) return st;
    if (ahead) EVP_CHECK(hipMemcpy(&c->h_ctl[0], c->ctl, sizeof(LoopCtl), hipMemcpyDeviceToHost));
    const LoopCtl &L = c->h_ctl[0];
    if (L.f.status) return L.f.status;
    // every iteration carried out swapped the stress buffers once; the
    // queued ones that returned at once did not
    if ((queued - (L.f.iter - iter0)) % 2) {
        double *t = c->sig;
        c->sig = c->sig_alt;
        c->sig_alt = t;
    }
    return EVPGPU_OK;
}

// E = Estart + P_compatible(E - Estart) in FP64 (uniform part unchanged)
int project_increment(evpgpu_ctx *c)
{
    int st = sub_launch(c, c->E, c->Estart, 6 * c->N);
    if (!st) st = fft_forward(c, c->E, 6, EVPGPU_FP64);
    if (!st) st = green_project(c);
    if (!st) st = fft_inverse_set(c, c->E, c->Estart, 1.0 / (double)c->N);
    return st;
}

int ready(const evpgpu_ctx *c)
{
    return c && c->have_phases && c->have_loading && c->have_ms && c->have_state;
}

} // namespace

// the Newton kernel's parameters of a context (for launches from other files)
NewtonParams context_newton_params(const evpgpu_ctx *c)
{
    return newton_params(c);
}

// ------------------------------------------------------------------- ABI --
extern "C" {

const char *evpgpu_strerror(int s)
{
    switch (s) {
    case EVPGPU_OK: return "ok";
    case EVPGPU_ERR_SINGULAR: return "singular 6x6 matrix in LU decomposition (35)";
    case EVPGPU_ERR_LINESEARCH: return "roundoff problem during line searching in NR (1203)";
    case EVPGPU_ERR_ARG: return "invalid argument";
    case EVPGPU_ERR_HIP: return "HIP runtime error";
    case EVPGPU_ERR_SIZE: return "unsupported grid size (each length 2..512)";
    case EVPGPU_ERR_NOMEM: return "device allocation failed";
    case EVPGPU_ERR_STATE: return "call out of order (phases, loading, microstructure, state)";
    case EVPGPU_ERR_NODEV: return "no HIP device";
    case EVPGPU_ERR_NONFINITE: return "non-finite stress or error norm in the stress update";
    case EVPGPU_ERR_PF_SUM: return "order parameters of a voxel sum below 0.5 (1385)";
    case EVPGPU_ERR_PF_CAP: return "order-parameter list beyond EVPGPU_PF_CAP entries";
    case EVPGPU_ERR_PF_GRAIN: return "grain index outside rho_avg";
    }
    return "unknown status";
}

int evpgpu_create(evpgpu_ctx **out, int nx, int ny, int nz, int precision)
{
    if (!out) return EVPGPU_ERR_ARG;
    *out = nullptr;
    if (!len_ok(nx) || !len_ok(ny) || !len_ok(nz)) return EVPGPU_ERR_SIZE;
    if (precision != EVPGPU_FP64 && precision != EVPGPU_MIXED) return EVPGPU_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return EVPGPU_ERR_NODEV;
    evpgpu_ctx *c = (evpgpu_ctx *)calloc(1, sizeof(evpgpu_ctx));
    EVP_CHECK(hipGetDevice(&c->device));
    c->nx = nx; c->ny = ny; c->nz = nz; c->nzh = nz / 2 + 1;
    c->N = (long long)nx * ny * nz;
    c->NK = (long long)nx * ny * c->nzh;
    c->precision = precision;
    int st = EVPGPU_OK;
    if (hipStreamCreate(&c->stream) != hipSuccess) st = EVPGPU_ERR_HIP;
    const long long N = c->N;
    // the Newton kernel runs 64-lane blocks and writes NRED partials per block;
    // the epilogue (256-lane blocks, NRED_EPI partials) needs less
    long long nblk = (N + 63) / 64;
    if (!st) st = dalloc(&c->sig, 6 * N);
    if (!st) st = dalloc(&c->sig_alt, 6 * N);
    if (!st) st = dalloc(&c->E, 6 * N);
    if (!st) st = dalloc(&c->W, 3 * N);
    if (!st) st = dalloc(&c->epsp, 6 * N);
    if (!st) st = dalloc(&c->edot, 6 * N);
    if (!st) st = dalloc(&c->tran, 9 * N);
    if (!st) st = dalloc(&c->crss, 24 * N);
    if (!st) st = dalloc(&c->gamacum, N);
    if (!st) st = dalloc(&c->gamdot, 12 * N);
    if (!st) st = dalloc(&c->phase, N);
    if (!st) st = dalloc(&c->grain, N);
    if (!st) st = dalloc(&c->orient, N);
    if (!st) st = dalloc((double2 **)&c->spec, 6 * c->NK);
    if (!st) st = dalloc(&c->acc, 3 * c->NK);
    if (!st) st = dalloc(&c->rwork, 3 * N);
    if (!st) st = dalloc(&c->sigA, 6 * N);
    if (!st) st = dalloc(&c->cA, 6 * N);
    if (!st) st = dalloc(&c->skA, 3 * c->NK);
    if (!st) st = dalloc(&c->Estart, 6 * N);
    c->anchor_every = 16;
    c->nr_tol = 1E-10;
    c->tolx = 1.E-7;
    c->newton_shortcut = 1;
    c->fused = 1;
    if (!st) st = dalloc(&c->partials, nblk * NRED);
    if (!st) st = dalloc(&c->red, NRED + 1);
    if (!st) st = dalloc(&c->red2, (long long)REDUCE_G * NRED);
    if (!st) st = dalloc(&c->d_status, 1);
    if (!st && hipHostMalloc((void **)&c->h_red, (NRED + 1) * sizeof(double), 0) != hipSuccess) st = EVPGPU_ERR_NOMEM;
    if (!st && hipHostMalloc((void **)&c->h_status, sizeof(int), 0) != hipSuccess) st = EVPGPU_ERR_NOMEM;
    if (!st && hipMalloc((void **)&c->ctl, sizeof(LoopCtl)) != hipSuccess) st = EVPGPU_ERR_NOMEM;
    if (!st && hipMalloc((void **)&c->ctl_k, sizeof(CtlConst)) != hipSuccess) st = EVPGPU_ERR_NOMEM;
    if (!st && hipHostMalloc((void **)&c->h_ctl, 3 * sizeof(LoopCtl), 0) != hipSuccess) st = EVPGPU_ERR_NOMEM;
    for (int b = 0; b < 2 && !st; b++)
        if (hipEventCreateWithFlags(&c->bev[b], hipEventDisableTiming) != hipSuccess) st = EVPGPU_ERR_HIP;
    c->loop_batch = 4;
    if (!st) st = fft_twiddles(c);
    if (!st) {
        // fields start at zero (gamdot of gas voxels is never written)
        st = fill_zero(c, c->gamdot, 12 * N);
        if (!st && hipMemsetAsync(c->d_status, 0, sizeof(int), c->stream) != hipSuccess) st = EVPGPU_ERR_HIP;
        if (!st && hipStreamSynchronize(c->stream) != hipSuccess) st = EVPGPU_ERR_HIP;
    }
    if (st) {
        evpgpu_destroy(c);
        return st;
    }
    *out = c;
    return EVPGPU_OK;
}

int evpgpu_destroy(evpgpu_ctx *c)
{
    if (!c) return EVPGPU_OK;
    adjoint_free(c);
    void *ptrs[] = {c->sig, c->sig_alt, c->E, c->W, c->epsp, c->edot, c->tran, c->crss, c->gamacum, c->gamdot,
                    c->phase, c->grain, c->orient, c->sgr, c->spec, c->acc, c->rwork,
                    c->sigA, c->cA, c->skA, c->Estart, c->partials, c->red, c->red2, c->d_status, c->ctl, c->ctl_k, c->tw64[0], c->tw64[1], c->tw64[2],
                    c->tw32[0], c->tw32[1], c->tw32[2], c->rho, c->dd_k, c->dd_tab};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++)
        if (ptrs[i]) (void)hipFree(ptrs[i]);
    for (int q = 0; q < EVP_NSLOT; q++)
        for (int k = 0; k < EVP_NTIMERS; k++)
            for (int e = 0; e < 2; e++)
                if (c->ev[q][k][e]) (void)hipEventDestroy(c->ev[q][k][e]);
    for (int b = 0; b < 2; b++)
        if (c->bev[b]) (void)hipEventDestroy(c->bev[b]);
    free(c->h_sgr);
    if (c->h_red) (void)hipHostFree(c->h_red);
    if (c->h_ctl) (void)hipHostFree(c->h_ctl);
    if (c->h_status) (void)hipHostFree(c->h_status);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    free(c);
    return EVPGPU_OK;
}

int evpgpu_set_timing(evpgpu_ctx *c, int on)
{
    if (!c) return EVPGPU_ERR_ARG;
    if (on && !c->ev[EVP_NSLOT - 1][EVP_NTIMERS - 1][1])
        for (int q = 0; q < EVP_NSLOT; q++)
            for (int k = 0; k < EVP_NTIMERS; k++)
                for (int e = 0; e < 2; e++)
                    if (!c->ev[q][k][e]) EVP_CHECK(hipEventCreate(&c->ev[q][k][e]));
    c->timing = on ? 1 : 0;
    c->timing_groups = (1u << EVP_NTIMERS) - 1;
    c->timing_every = 1;
    c->tcount = 0;
    c->tskip = 0;
    c->tslot = 0;
    memset(c->tmask, 0, sizeof(c->tmask));
    memset(c->t_ms, 0, sizeof(c->t_ms));
    memset(c->t_n, 0, sizeof(c->t_n));
    return EVPGPU_OK;
}

int evpgpu_synchronize(evpgpu_ctx *c)
{
    if (!c) return EVPGPU_ERR_ARG;
    EVP_CHECK(hipStreamSynchronize(c->stream));
    EVP_CHECK(hipDeviceSynchronize());
    return EVPGPU_OK;
}

int evpgpu_get_timing(evpgpu_ctx *c, int which, double *ms, long long *launches)
{
    if (!c || which < 0 || which >= EVP_NTIMERS) return EVPGPU_ERR_ARG;
    if (ms) *ms = c->t_ms[which];
    if (launches) *launches = c->t_n[which];
    return EVPGPU_OK;
}

int evpgpu_set_timing_groups(evpgpu_ctx *c, unsigned groups, int every)
{
    if (!c || groups >= (1u << EVP_NTIMERS) || every < 1) return EVPGPU_ERR_ARG;
    c->timing_groups = groups;
    c->timing_every = every;
    c->tcount = 0;
    return EVPGPU_OK;
}

int evpgpu_set_loop_batch(evpgpu_ctx *c, int batch)
{
    if (!c || batch < 0 || batch > EVP_NSLOT / 2) return EVPGPU_ERR_ARG;
    c->loop_batch = batch;
    return EVPGPU_OK;
}

int evpgpu_set_fused_transforms(evpgpu_ctx *c, int on)
{
    if (!c) return EVPGPU_ERR_ARG;
    c->fused = on ? 1 : 0;
    return EVPGPU_OK;
}

int evpgpu_set_newton_shortcut(evpgpu_ctx *c, int on)
{
    if (!c) return EVPGPU_ERR_ARG;
    c->newton_shortcut = on ? 1 : 0;
    return EVPGPU_OK;
}

int evpgpu_set_newton_tolerances(evpgpu_ctx *c, double tol, double tolx)
{
    if (!c || !(tol > 0.0) || !(tolx > 0.0)) return EVPGPU_ERR_ARG;
    c->nr_tol = tol;
    c->tolx = tolx;
    return EVPGPU_OK;
}

int evpgpu_set_anchor_interval(evpgpu_ctx *c, int iterations)
{
    if (!c || iterations < 1) return EVPGPU_ERR_ARG;
    c->anchor_every = iterations;
    return EVPGPU_OK;
}

int evpgpu_set_precision(evpgpu_ctx *c, int precision)
{
    if (!c || (precision != EVPGPU_FP64 && precision != EVPGPU_MIXED)) return EVPGPU_ERR_ARG;
    c->precision = precision;
    return EVPGPU_OK;
}

int evpgpu_set_phases(evpgpu_ctx *c, const evpgpu_phase *ph, int n)
{
    if (!c || !ph || n < 1 || n > 2) return EVPGPU_ERR_ARG;
    for (int p = 0; p < n; p++)
        if (ph[p].nsys < 0 || ph[p].nsys > EVPGPU_NSYSMX) return EVPGPU_ERR_ARG;
    memcpy(c->ph, ph, sizeof(evpgpu_phase) * n);
    c->nphases = n;
    c->have_phases = 1;
    return EVPGPU_OK;
}

int evpgpu_set_loading(evpgpu_ctx *c, const evpgpu_loading *ld)
{
    if (!c || !ld) return EVPGPU_ERR_ARG;
    c->ld = *ld;
    // Udot = VoigtToFull(VelGrad_BC) (src/kinematics.c:5-12), Scauchy likewise
    const double *v = ld->velgrad_bc, *s = ld->stress_bc;
    double U[9] = {v[0], v[5], v[4], v[5], v[1], v[3], v[4], v[3], v[2]};
    double S[9] = {s[0], s[5], s[4], s[5], s[1], s[3], s[4], s[3], s[2]};
    memcpy(c->udot, U, sizeof(U));
    memcpy(c->scauchy, S, sizeof(S));
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            c->udot_s[a * 3 + b] = (U[a * 3 + b] + U[b * 3 + a]) / 2.;
            c->udot_a[a * 3 + b] = (U[a * 3 + b] - U[b * 3 + a]) / 2.;
        }
    c->creep_flag = 0;
    for (int i = 0; i < 6; i++) c->creep_flag += ld->stress_bc_flag[i];
    c->creep_flag /= 6;
    c->have_loading = 1;
    update_lam_min(c);
    return EVPGPU_OK;
}

int evpgpu_upload_microstructure(evpgpu_ctx *c, const int *grain, const int *phase, const int *orient,
                                 int n_orient, const double *sgr)
{
    if (!c || !grain || !phase || !orient || n_orient < 1 || !sgr) return EVPGPU_ERR_ARG;
    for (long long v = 0; v < c->N; v++)
        if (orient[v] >= n_orient || (orient[v] < 0 && phase[v] >= 1 && phase[v] <= c->nphases &&
                                      !c->ph[phase[v] - 1].is_gas))
            return EVPGPU_ERR_ARG;
    // gas voxels carry orientation -1; the kernel never reads their table row,
    // but keep the index valid
    int *o2 = (int *)malloc(sizeof(int) * c->N);
    for (long long v = 0; v < c->N; v++) o2[v] = orient[v] < 0 ? 0 : orient[v];
    if (c->sgr) (void)hipFree(c->sgr);
    int st = dalloc(&c->sgr, 36LL * n_orient);
    if (!st && hipMemcpy(c->sgr, sgr, sizeof(double) * 36 * n_orient, hipMemcpyHostToDevice) != hipSuccess) st = EVPGPU_ERR_HIP;
    if (!st && hipMemcpy(c->grain, grain, sizeof(int) * c->N, hipMemcpyHostToDevice) != hipSuccess) st = EVPGPU_ERR_HIP;
    if (!st && hipMemcpy(c->phase, phase, sizeof(int) * c->N, hipMemcpyHostToDevice) != hipSuccess) st = EVPGPU_ERR_HIP;
    if (!st && hipMemcpy(c->orient, o2, sizeof(int) * c->N, hipMemcpyHostToDevice) != hipSuccess) st = EVPGPU_ERR_HIP;
    free(o2);
    if (st) return st;
    c->n_orient = n_orient;
    free(c->h_sgr);
    c->h_sgr = (double *)malloc(sizeof(double) * 36 * n_orient);
    memcpy(c->h_sgr, sgr, sizeof(double) * 36 * n_orient);
    update_lam_min(c);
    c->have_ms = 1;
    return EVPGPU_OK;
}

int evpgpu_upload_state(evpgpu_ctx *c, const double *sig6, const double *dg9, const double *eps6,
                        const double *edot6, const double *tran9, const double *crss24,
                        const double *gamacum)
{
    if (!c || !sig6 || !dg9 || !eps6 || !edot6 || !tran9 || !crss24 || !gamacum) return EVPGPU_ERR_ARG;
    if (!c->have_phases || !c->have_ms) return EVPGPU_ERR_STATE;
    const long long N = c->N;
    double *E = (double *)malloc(sizeof(double) * 9 * N);
    double *W = E + 6 * N;
    for (long long v = 0; v < N; v++) {
        const double *d = dg9;
        double g[9];
        for (int k = 0; k < 9; k++) g[k] = d[k * N + v];
        E[0 * N + v] = g[0];
        E[1 * N + v] = g[4];
        E[2 * N + v] = g[8];
        E[3 * N + v] = (g[5] + g[7]) / 2.0;
        E[4 * N + v] = (g[2] + g[6]) / 2.0;
        E[5 * N + v] = (g[1] + g[3]) / 2.0;
        W[0 * N + v] = (g[5] - g[7]) / 2.0;
        W[1 * N + v] = (g[2] - g[6]) / 2.0;
        W[2 * N + v] = (g[1] - g[3]) / 2.0;
    }
    int sym = 1;
    for (long long i = 0; i < 12 * N && sym; i++) {
        long long s = i / N, v = i % N;
        if (crss24[(2 * s) * N + v] != crss24[(2 * s + 1) * N + v]) sym = 0;
    }
    c->crss_sym = sym;
    int st = EVPGPU_OK;
#define H2D(dst, src, n) \
    if (!st && hipMemcpy(dst, src, sizeof(double) * (n), hipMemcpyHostToDevice) != hipSuccess) st = EVPGPU_ERR_HIP
    H2D(c->sig, sig6, 6 * N);
    H2D(c->E, E, 6 * N);
    H2D(c->W, W, 3 * N);
    H2D(c->epsp, eps6, 6 * N);
    H2D(c->edot, edot6, 6 * N);
    H2D(c->tran, tran9, 9 * N);
    H2D(c->crss, crss24, 24 * N);
    H2D(c->gamacum, gamacum, N);
#undef H2D
    free(E);
    if (st) return st;
    // (the Newton kernel derives the Schmid tensors from the current
    // orientations, so they are initialised at setup as well, Q12)
    if (!st && hipStreamSynchronize(c->stream) != hipSuccess) st = EVPGPU_ERR_HIP;
    if (!st) c->have_state = 1;
    return st;
}

int evpgpu_download_state(evpgpu_ctx *c, double *sig6, double *dg9, double *eps6, double *edot6,
                          double *tran9, double *crss24, double *gamacum, double *gamdot12)
{
    if (!c) return EVPGPU_ERR_ARG;
    const long long N = c->N;
    EVP_CHECK(hipStreamSynchronize(c->stream));
#define D2H(dst, src, n) \
    if (dst) EVP_CHECK(hipMemcpy(dst, src, sizeof(double) * (n), hipMemcpyDeviceToHost))
    D2H(sig6, c->sig, 6 * N);
    D2H(eps6, c->epsp, 6 * N);
    D2H(edot6, c->edot, 6 * N);
    D2H(tran9, c->tran, 9 * N);
    D2H(crss24, c->crss, 24 * N);
    D2H(gamacum, c->gamacum, N);
    D2H(gamdot12, c->gamdot, 12 * N);
#undef D2H
    if (dg9) {
        double *E = (double *)malloc(sizeof(double) * 9 * N), *W = E + 6 * N;
        if (hipMemcpy(E, c->E, sizeof(double) * 6 * N, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(W, c->W, sizeof(double) * 3 * N, hipMemcpyDeviceToHost) != hipSuccess) {
            free(E);
            return EVPGPU_ERR_HIP;
        }
        for (long long v = 0; v < N; v++) {
            double e[6], w[3];
            for (int k = 0; k < 6; k++) e[k] = E[k * N + v];
            for (int k = 0; k < 3; k++) w[k] = W[k * N + v];
            double g[9] = {e[0], e[5] + w[2], e[4] + w[1], e[5] - w[2], e[1],
                           e[3] + w[0], e[4] - w[1], e[3] - w[0], e[2]};
            for (int k = 0; k < 9; k++) dg9[k * N + v] = g[k];
        }
        free(E);
    }
    return EVPGPU_OK;
}

int evpgpu_set_macro(evpgpu_ctx *c, const evpgpu_macro *m)
{
    if (!c || !m) return EVPGPU_ERR_ARG;
    c->m = *m;
    memcpy(c->disgrad_avg_actual, m->disgrad_avg_t, sizeof(c->disgrad_avg_actual));
    return EVPGPU_OK;
}

int evpgpu_get_macro(evpgpu_ctx *c, evpgpu_macro *m)
{
    if (!c || !m) return EVPGPU_ERR_ARG;
    *m = c->m;
    return EVPGPU_OK;
}

// One loading increment.  forced > 0 runs exactly that many iterations;
// rep64 >= 0 replays a recorded increment: rep_mix mixed iterations (mixed
// precision only), the projection, then rep64 FP64 iterations -- the same
// kernel sequence as the run that recorded it, so the same result bit for bit.
int increment_core(evpgpu_ctx *c, int forced, int rep_mix, int rep64, evpgpu_stats *stats)
{
    if (!ready(c)) return EVPGPU_ERR_STATE;
    evpgpu_stats S;
    memset(&S, 0, sizeof(S));
    const double dt = c->ld.dt;
    evpgpu_macro &m = c->m;
    if (c->creep_flag == 1 && m.e_vm > 0.3) {
        S.stopped = 1;
        if (stats) *stats = S;
        return EVPGPU_OK;
    }
    // prologue: DisGrad += Udot*dt (creep: DisGrad = DisGradAvg_actual)
    double pro[9];
    const double *src = c->creep_flag == 1 ? c->disgrad_avg_actual : nullptr;
    double U[9];
    for (int i = 0; i < 9; i++) U[i] = src ? src[i] : c->udot[i] * dt;
    pro[0] = U[0]; pro[1] = U[4]; pro[2] = U[8];
    pro[3] = (U[5] + U[7]) / 2.0; pro[4] = (U[2] + U[6]) / 2.0; pro[5] = (U[1] + U[3]) / 2.0;
    pro[6] = (U[5] - U[7]) / 2.0; pro[7] = (U[2] - U[6]) / 2.0; pro[8] = (U[1] - U[3]) / 2.0;
    tmark(c, 4, 0);
    int st = prologue_launch(c, pro, c->creep_flag == 1, nullptr);
    if (st) return st;
    if (c->dd_on) {
        // the rate-law constants of this increment's densities; the first
        // Newton launch reads the first call's passing stress
        st = dd_prologue_launch(c, dd_params(c));
        if (st) return st;
        c->dd_first = 1;
    }
    // update_schmid (src/evolution.c:1334-1336) happens inside the Newton
    // kernel, from the orientations the previous epilogue left
    st = fill_zero(c, (double *)c->acc, 6 * c->NK);
    // a failed Newton launch of an earlier increment leaves its status word
    // set; every increment starts from a clear one
    if (!st) st = fill_zero_int(c, c->d_status, 1);
    if (st) return st;
    tmark(c, 4, 1);
    EVP_CHECK(hipStreamSynchronize(c->stream));
    tcollect(c);
    // the control block at the start of the loop: no macro strain correction
    // yet, error norms above Err; the macro stresses stay the last
    // increment's when the loop runs no iteration
    LoopCtl init;
    memset(&init, 0, sizeof(init));
    init.tol = c->ld.err_tol;
    init.e_vm = m.e_vm;
    init.err_e = 2. * c->ld.err_tol;
    init.err_s = 2. * c->ld.err_tol;
    memcpy(init.sig_avg, m.sig_avg, sizeof(init.sig_avg));
    memcpy(init.sig_avg1, m.sig_avg1, sizeof(init.sig_avg1));
    init.s_vm = m.s_vm;
    init.s_vm1 = m.s_vm1;
    c->h_ctl[0] = init;
    const CtlConst K = ctl_const(c);
    st = ctl_init_launch(c, init, K);
    if (st) return st;
    double t0 = now_s();
    const bool mixed = c->precision == EVPGPU_MIXED;
    // the strain at the start of the loop: the mixed loop's projection and
    // the antisymmetric increment at its end use it
    EVP_CHECK(hipMemcpyAsync(c->Estart, c->E, 6 * c->N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    int imix = 0, it_mixed = 0;
    // the loop's phases (limits count every iteration of the increment)
    auto mixed_phase = [&](int limit, int conv) {
        int r = run_phase(c, EVPGPU_MIXED, limit, conv, 0, K, &imix);
        it_mixed = c->h_ctl[0].f.iter;
        if (!r) r = project_increment(c);
        return r;
    };
    auto fp64_phase = [&](int limit, int conv, int first) {
        return run_phase(c, EVPGPU_FP64, limit, conv, first, K, &imix);
    };
    if (rep64 >= 0) {
        if (mixed) st = mixed_phase(rep_mix, 0);
        if (!st) st = fp64_phase(c->h_ctl[0].f.iter + rep64, 0, 0);
    } else if (forced > 0) {
        if (mixed) st = mixed_phase(forced - 1, 0);
        if (!st) st = fp64_phase(forced, 0, 0);
    } else if (mixed) {
        // the FP64 sweep that closes the mixed loop counts against IterMax:
        // the mixed iterations stop one short of it; then FP64 correction
        // sweeps, at least one, until converged
        st = mixed_phase(c->ld.iter_max - 1, 1);
        if (!st) st = fp64_phase(c->ld.iter_max, 1, 1);
    } else {
        st = fp64_phase(c->ld.iter_max, 1, 0);
    }
    if (st) return st;
    const LoopCtl &L = c->h_ctl[0];
    const int iter = L.f.iter, it64 = iter - it_mixed;
    m.err_e = L.err_e;
    m.err_s = L.err_s;
    memcpy(m.sig_avg, L.sig_avg, sizeof(m.sig_avg));
    memcpy(m.sig_avg1, L.sig_avg1, sizeof(m.sig_avg1));
    m.s_vm = L.s_vm;
    m.s_vm1 = L.s_vm1;
    memcpy(c->dD, L.dD, sizeof(c->dD));
    memcpy(c->dD_acum, L.dD_acum, sizeof(c->dD_acum));
    memcpy(c->dD_applied, L.dD_applied, sizeof(c->dD_applied));
    // The loop's Newton launches may end an iteration whose step the
    // reference rejects before evaluating its trial point; gamdot comes from
    // that trial (Q5) and is read only by the epilogue, so the last launch is
    // repeated with every iteration carried out (same lambda, same sigma,
    // edot and norms; gamdot as the reference leaves it).
    {
        NewtonParams P = newton_params(c);
        P.dd_first = c->dd_on && iter == 1;
        int nblk = 0;
        hipEvent_t e0, e1;
        tkernel(c, 0, &e0, &e1);
        st = newton_launch(c, P, c->sig_alt, c->sig, c->E, c->epsp, c->edot, c->tran,
                           c->dd_on ? c->dd_k : c->crss, c->gamdot, c->phase, c->orient, c->sgr, c->N,
                           c->partials, &nblk, c->d_status, nullptr, nullptr, 0, e0, e1);
        if (st) return st;
        EVP_CHECK(hipMemcpyAsync(c->h_status, c->d_status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        EVP_CHECK(hipStreamSynchronize(c->stream));
        tcollect(c);
        if (*c->h_status) return *c->h_status;
    }

    S.seconds = now_s() - t0;
    S.err_s = m.err_s;
    S.err_e = m.err_e;
    S.iters = iter;
    S.fp64_iters = it64;
    S.newton_iters = L.nit;
    S.ls_trials = L.ntr;

    // epilogue (src/evolution.c:1469-1547)
    // antisymmetric displacement-gradient increment: uniform part + IFFT of the
    // accumulated antisymmetric Green output
    double m3[3];
    {
        const double *a = c->udot_a, *d = c->dD_applied;
        m3[0] = a[5] * dt + (d[5] - d[7]) / 2.0;
        m3[1] = a[2] * dt + (d[2] - d[6]) / 2.0;
        m3[2] = a[1] * dt + (d[1] - d[3]) / 2.0;
        if (c->creep_flag == 1) m3[0] = m3[1] = m3[2] = 0.0;
    }
    tmark(c, 4, 0);
    if (c->dd_on) {
        // gradient of the shear rates and DislocationEvolution, with the
        // orientations of the loop (before update_orient)
        st = dd_evolve_launch(c, dd_params(c));
        if (st) return st;
    }
    // the antisymmetric part of the loop's Green outputs: at every wave vector
    // but the split ones it follows from the strain change E - Estart (its
    // spectrum is the sum of the outputs' symmetric parts); the split ones
    // were accumulated in the loop
    st = rsub_launch(c, c->E, c->Estart, 6 * c->N);
    if (!st) st = fft_forward_to(c, c->Estart, 6, (double2 *)c->spec);
    if (!st) st = green_skew_fill(c);
    if (!st) st = fft_inverse_real(c, c->acc, 3, c->rwork, 1.0 / (double)c->N);
    if (st) return st;
    st = antisym_add_launch(c, m3);
    if (st) return st;
    double actual[9], vg[9];
    for (int i = 0; i < 9; i++) {
        actual[i] = m.disgrad_avg[i] + c->dD_acum[i];
        vg[i] = (actual[i] - m.disgrad_avg_t[i]) / dt;
    }
    memcpy(c->disgrad_avg_actual, actual, sizeof(actual));
    m.e_vm = vonmises(actual);
    m.d_vm = vonmises(vg);
    m.time_tot += dt;
    memcpy(m.disgrad_avg_t, actual, sizeof(actual));
    EpilogueParams E;
    memset(&E, 0, sizeof(E));
    E.dt = dt;
    memcpy(E.udot_a, c->udot_a, sizeof(E.udot_a));
    E.update_orient = (c->ld.update_flag == 1 && m.istep > 1);
    // the dislocation-density build does not call harden (src/evolution.c:1571-1576)
    E.harden = (c->ld.hard_flag == 1 && m.istep > 1) && !c->dd_on;
    E.wgt = 1.0 / (double)c->N;
    for (int p = 0; p < 2; p++) {
        E.is_gas[p] = p < c->nphases ? c->ph[p].is_gas : 1;
        if (p >= c->nphases) continue;
        const evpgpu_phase &ph = c->ph[p];
        E.nsys[p] = ph.nsys;
        memcpy(E.hard[p], ph.hard, sizeof(E.hard[p]));
        memcpy(E.tau[p], ph.tau, sizeof(E.tau[p]));
        memcpy(E.thet[p], ph.thet, sizeof(E.thet[p]));
        memcpy(E.bn[p], ph.bn, sizeof(E.bn[p]));
        memcpy(E.bb[p], ph.bb, sizeof(E.bb[p]));
    }
    int nb = 0;
    st = epilogue_launch(c, E, c->partials, &nb);
    if (st) return st;
    st = reduce_partials(c, c->partials, nb, NRED_EPI, c->red);
    if (st) return st;
    tmark(c, 4, 1);
    EVP_CHECK(hipMemcpyAsync(c->h_red, c->red, NRED_EPI * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    EVP_CHECK(hipStreamSynchronize(c->stream));
    tcollect(c);
    sym6_to_full(c->h_red, m.eps_avg);
    sym6_to_full(c->h_red + 6, m.edot_avg);
    double ev = 0.0, dv = 0.0;
    for (int i = 0; i < 9; i++) {
        ev += m.eps_avg[i] * m.eps_avg[i];
        dv += m.edot_avg[i] * m.edot_avg[i];
    }
    m.evmp = sqrt(2. / 3. * ev);
    m.dvmp = sqrt(2. / 3. * dv);
    for (int i = 0; i < 9; i++) m.disgrad_avg[i] = m.disgrad_avg_t[i] + c->udot[i] * dt;
    m.istep++;
    if (stats) *stats = S;
    return EVPGPU_OK;
}

int evpgpu_increment(evpgpu_ctx *c, int forced, evpgpu_stats *stats)
{
    return increment_core(c, forced, 0, -1, stats);
}

// ------------------------------------------------------ building blocks --
int evpgpu_copy_probe(evpgpu_ctx *c, long long n, int reps, double *gbs)
{
    if (!c || n < 2 || (n & 1) || reps < 1 || !gbs) return EVPGPU_ERR_ARG;
    double *a = nullptr, *b = nullptr;
    int st = dalloc(&a, n);
    if (!st) st = dalloc(&b, n);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!st && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) st = EVPGPU_ERR_HIP;
    if (!st) st = fill_zero(c, a, n);
    if (!st) st = copy_launch(c, a, b, n); // warm
    if (!st && hipEventRecord(e0, c->stream) != hipSuccess) st = EVPGPU_ERR_HIP;
    for (int r = 0; r < reps && !st; r++) st = copy_launch(c, a, b, n);
    if (!st && hipEventRecord(e1, c->stream) != hipSuccess) st = EVPGPU_ERR_HIP;
    if (!st && hipEventSynchronize(e1) != hipSuccess) st = EVPGPU_ERR_HIP;
    float ms = 0.f;
    if (!st && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = EVPGPU_ERR_HIP;
    if (!st) *gbs = 16.0 * (double)n * reps / (1e-3 * ms) / 1e9;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return st;
}

int evpgpu_fft_r2c_host(evpgpu_ctx *c, const double *in, double *out)
{
    if (!c || !in || !out) return EVPGPU_ERR_ARG;
    EVP_CHECK(hipMemcpy(c->rwork, in, sizeof(double) * c->N, hipMemcpyHostToDevice));
    int st = fft_forward(c, c->rwork, 1, c->precision);
    if (st) return st;
    EVP_CHECK(hipStreamSynchronize(c->stream));
    if (c->precision == EVPGPU_MIXED) {
        float *tmp = (float *)malloc(sizeof(float) * 2 * c->NK);
        EVP_CHECK(hipMemcpy(tmp, c->spec, sizeof(float) * 2 * c->NK, hipMemcpyDeviceToHost));
        for (long long i = 0; i < 2 * c->NK; i++) out[i] = tmp[i];
        free(tmp);
    } else {
        EVP_CHECK(hipMemcpy(out, c->spec, sizeof(double) * 2 * c->NK, hipMemcpyDeviceToHost));
    }
    return EVPGPU_OK;
}

int evpgpu_fft_c2r_host(evpgpu_ctx *c, const double *in, double *out)
{
    if (!c || !in || !out) return EVPGPU_ERR_ARG;
    if (c->precision == EVPGPU_MIXED) {
        float *tmp = (float *)malloc(sizeof(float) * 2 * c->NK);
        for (long long i = 0; i < 2 * c->NK; i++) tmp[i] = (float)in[i];
        EVP_CHECK(hipMemcpy(c->spec, tmp, sizeof(float) * 2 * c->NK, hipMemcpyHostToDevice));
        free(tmp);
    } else {
        EVP_CHECK(hipMemcpy(c->spec, in, sizeof(double) * 2 * c->NK, hipMemcpyHostToDevice));
    }
    int st = fill_zero(c, c->rwork, c->N);
    if (!st) st = fft_inverse_update(c, 1, c->precision, c->rwork, nullptr, 1.0);
    if (st) return st;
    EVP_CHECK(hipStreamSynchronize(c->stream));
    EVP_CHECK(hipMemcpy(out, c->rwork, sizeof(double) * c->N, hipMemcpyDeviceToHost));
    return EVPGPU_OK;
}

int evpgpu_green_host(evpgpu_ctx *c, const double *in6, double *out9)
{
    if (!c || !in6 || !out9 || !c->have_loading) return EVPGPU_ERR_ARG;
    double2 *o9 = nullptr;
    int st = dalloc(&o9, 9 * c->NK);
    if (st) return st;
    if (hipMemcpy(c->spec, in6, sizeof(double2) * 6 * c->NK, hipMemcpyHostToDevice) != hipSuccess) st = EVPGPU_ERR_HIP;
    if (!st) st = green_full9(c, (double2 *)c->spec, o9);
    if (!st && hipStreamSynchronize(c->stream) != hipSuccess) st = EVPGPU_ERR_HIP;
    if (!st && hipMemcpy(out9, o9, sizeof(double2) * 9 * c->NK, hipMemcpyDeviceToHost) != hipSuccess) st = EVPGPU_ERR_HIP;
    (void)hipFree(o9);
    return st;
}

int evpgpu_newton_host(evpgpu_ctx *c, int n, const int *phase, const int *orient, const double *sgr,
                       int n_orient, double *sig6, const double *strain6, const double *eps6,
                       const double *tranmat9, const double *crss24, double *edot6, double *gamdot12,
                       double *err2, int *niter)
{
    if (!c || n < 1 || !phase || !orient || !sgr || !sig6 || !strain6 || !eps6 || !tranmat9 || !crss24 ||
        !edot6 || !gamdot12 || !err2 || !niter || !c->have_phases || !c->have_loading)
        return EVPGPU_ERR_ARG;
    for (int v = 0; v < n; v++)
        if (orient[v] < 0 || orient[v] >= n_orient || phase[v] < 1 || phase[v] > c->nphases) return EVPGPU_ERR_ARG;
    int *dph = nullptr, *dor = nullptr, *dni = nullptr, *dst = nullptr;
    double *dsg = nullptr, *dsig = nullptr, *dstr = nullptr, *deps = nullptr, *dsc = nullptr,
           *dcr = nullptr, *ded = nullptr, *dgd = nullptr, *der = nullptr, *dpart = nullptr;
    int st = EVPGPU_OK;
    if (!st) st = dalloc(&dph, n);
    if (!st) st = dalloc(&dor, n);
    if (!st) st = dalloc(&dni, n);
    if (!st) st = dalloc(&dst, 1);
    if (!st) st = dalloc(&dsg, 36LL * n_orient);
    if (!st) st = dalloc(&dsig, 6LL * n);
    if (!st) st = dalloc(&dstr, 6LL * n);
    if (!st) st = dalloc(&deps, 6LL * n);
    if (!st) st = dalloc(&dsc, 9LL * n);
    if (!st) st = dalloc(&dcr, 24LL * n);
    if (!st) st = dalloc(&ded, 6LL * n);
    if (!st) st = dalloc(&dgd, 12LL * n);
    if (!st) st = dalloc(&der, 2LL * n);
    if (!st) st = dalloc(&dpart, (long long)((n + 63) / 64) * NRED);
#define UP(d, s, cnt, T) \
    if (!st && hipMemcpy(d, s, sizeof(T) * (cnt), hipMemcpyHostToDevice) != hipSuccess) st = EVPGPU_ERR_HIP
    UP(dph, phase, n, int);
    UP(dor, orient, n, int);
    UP(dsg, sgr, 36LL * n_orient, double);
    UP(dsig, sig6, 6LL * n, double);
    UP(dstr, strain6, 6LL * n, double);
    UP(deps, eps6, 6LL * n, double);
    UP(dsc, tranmat9, 9LL * n, double);
    UP(dcr, crss24, 24LL * n, double);
    UP(dgd, gamdot12, 12LL * n, double);
#undef UP
    if (!st && hipMemset(dst, 0, sizeof(int)) != hipSuccess) st = EVPGPU_ERR_HIP;
    if (!st) {
        NewtonParams P = newton_params(c);
        P.crss_sym = 0;
        P.dd = 0; // the power-law update on the caller's CRSS
        int nb = 0;
        st = newton_launch(c, P, dsig, dsig, dstr, deps, ded, dsc, dcr, dgd, dph, dor, dsg, n, dpart, &nb, dst,
                           der, dni, 0);
        if (!st && hipStreamSynchronize(c->stream) != hipSuccess) st = EVPGPU_ERR_HIP;
    }
    int kst = 0;
#define DN(h, d, cnt, T) \
    if (!st && hipMemcpy(h, d, sizeof(T) * (cnt), hipMemcpyDeviceToHost) != hipSuccess) st = EVPGPU_ERR_HIP
    DN(&kst, dst, 1, int);
    DN(sig6, dsig, 6LL * n, double);
    DN(edot6, ded, 6LL * n, double);
    DN(gamdot12, dgd, 12LL * n, double);
    DN(err2, der, 2LL * n, double);
    DN(niter, dni, n, int);
#undef DN
    void *ptrs[] = {dph, dor, dni, dst, dsg, dsig, dstr, deps, dsc, dcr, ded, dgd, der, dpart};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++)
        if (ptrs[i]) (void)hipFree(ptrs[i]);
    if (!st && kst) st = kst;
    return st;
}

} // extern "C"
