// evpgpu_impl.h -- internal state and kernel launchers of libevpgpu (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>

#include "../../include/evpgpu.h"
#include "loopctl_step.h"

#define EVP_CHECK(x)                                                   \
    do {                                                               \
        hipError_t e_ = (x);                                           \
        if (e_ != hipSuccess) return EVPGPU_ERR_HIP;                   \
    } while (0)


// ---------------------------------------------------------------- params --
// Per-launch constants of the per-voxel stress update (passed by value: they
// land in SGPRs / the kernel-argument segment, uniform across the wave).
// Packed upper triangle of a symmetric 6x6 matrix: PK(i, j), i <= j.
__host__ __device__ constexpr int PK(int i, int j)
{
    return i <= j ? i * 6 - i * (i - 1) / 2 + (j - i) : j * 6 - j * (j - 1) / 2 + (i - j);
}

struct NewtonParams {
    double c0p[21];  // C0 (B-basis), packed upper triangle (PK)
    double dt;
    double nrs[2][EVPGPU_NSYSMX];
    double gam0[2][EVPGPU_NSYSMX];
    int nsys[2];
    int is_gas[2];
    int int_n[2];    // integer rate exponent of the phase, 0 if not integral
    int crss_sym;    // 1: forward and backward CRSS are equal in every voxel
    double bn[2][EVPGPU_NSYSMX][3], bb[2][EVPGPU_NSYSMX][3]; // crystal-axis normals, directions
    double s0p[21];  // S0 = C0^-1 (B-basis), packed upper triangle: the Newton step is
                     // solved in S0-scaled form
    double kmin_tolx2; // (0.99 TOLX min over orientations of lambda_min(S0 + S))^2
    double nr_tol;   // Newton tolerance on the relative step (reference 1e-10)
    double tolx;     // line-search TOLX (reference 1e-7)
    int dd;          // dislocation-density law: the crss argument holds the
                     // per-system passing stress, 1/(tau_0 + tau_cut), gam_o
                     // and the first call's passing stress, each 12N
    int dd_first;    // this launch makes each voxel's first rate call of the
                     // increment (the passing stress then uses the previous
                     // call's mobile density)
    const int *skip; // device flag: the launch does nothing when it is set
                     // (an iteration queued past the end of the Green loop)
};

struct GreenParams {
    // A = C0 : (n x n) is symmetric: A_ik (ik = 00 11 22 12 02 01) =
    // sum_p ga[ik*6 + p] n_j n_l over the pairs p = jl (00 11 22 12 02 01),
    // ga holding C0_ijkj for j = l and C0_ijkl + C0_ilkj for j < l
    double ga[36];
    double s0[81];
    int elast_bc;
};

struct EpilogueParams {
    double dt;
    double udot_a[9];
    int update_orient;
    int harden;
    int nsys[2];
    int is_gas[2];
    double wgt;
    double hard[2][EVPGPU_NSYSMX][EVPGPU_NSYSMX];
    double tau[2][EVPGPU_NSYSMX][3];
    double thet[2][EVPGPU_NSYSMX][2];
    double bn[2][EVPGPU_NSYSMX][3];
    double bb[2][EVPGPU_NSYSMX][3];
    double nrs[2][EVPGPU_NSYSMX];
    double gam0[2][EVPGPU_NSYSMX];
    int int_n[2];
    int crss_sym;
};

// number of doubles each Newton block writes to the partials array:
// err_s, err_e, newton iterations, sum sig (6), sum sig over phase 1 (6),
// line-search trials
// timer slots: 0 Newton, 1 forward FFT, 2 Green, 3 inverse FFT + strain update,
// 4 increment prologue/epilogue, 5 anchor refresh of the mixed-precision loop
#define EVP_NTIMERS 6
// timer slots per timer: one per Green iteration queued and not yet collected
#define EVP_NSLOT 8

#define NRED 16
// epilogue partials: sum eps*wgt (6), sum edot*wgt (6)
#define NRED_EPI 12
// first-stage workgroups of the partial-sum reduction
#define REDUCE_G 512


struct evpgpu_ctx {
    int device;
    int nx, ny, nz, nzh;
    long long N, NK;
    int precision;
    hipStream_t stream;

    int nphases;
    evpgpu_phase ph[2];
    evpgpu_loading ld;
    int have_phases, have_loading, have_ms, have_state;
    int crss_sym;

    // per-voxel fields (FP64, SoA component-major)
    double *sig;      // 6N  current stress
    double *sig_alt;  // 6N  the other stress buffer: the Newton reads lambda from
                      //     one and writes sigma to the other
    double *E;        // 6N  symmetric part of DisGrad
    double *W;        // 3N  antisymmetric part of DisGrad: (23, 13, 12)
    double *epsp;     // 6N  plastic strain (Eps)
    double *edot;     // 6N
    double *tran;     // 9N  TranMat_xt2sa
    double *crss;     // 24N (s*2+d)
    double *gamacum;  // N
    double *gamdot;   // 12N
    int *phase, *grain, *orient;
    double *sgr;      // n_orient x 36: LU_inv_66(C_gr)
    double *h_sgr;    // host copy
    double lam_min;   // min over orientations of lambda_min(S0 + S), < 0 when stale
    int n_orient;

    // spectral buffers
    void *spec;       // 6*NK complex (double2 or float2)
    double2 *acc;     // 3*NK complex: antisymmetric e-hat summed over the increment
    double *rwork;    // 3N real scratch
    // mixed-precision loop: FP32 anchor stress, its FP32 real-space Green
    // correction and antisymmetric spectrum; FP64 strain at the loop start
    float *sigA;      // 6N
    float *cA;        // 6N
    float2 *skA;      // 3*NK
    double *Estart;   // 6N
    int anchor_every; // iterations between FP64 anchor transforms
    double nr_tol, tolx; // Newton tolerances (evpgpu_set_newton_tolerances)
    int newton_shortcut; // end certainly-rejected Newton iterations early (default 1)
    int fused;           // one-pass-per-axis-pair FFT/Green kernels where they fit (default 1)
    double2 *tw64[3];
    float2 *tw32[3];

    // reductions
    double *partials; // max(nblk) * NRED
    double *red;      // NRED (device)
    double *red2;     // REDUCE_G * NRED, first stage of the reduction
    double *h_red;    // NRED pinned host
    int *d_status;
    int *h_status;

    // device-side loop control (loopctl.hip)
    LoopCtl *ctl;     // device
    CtlConst *ctl_k;  // device: the increment's constants
    LoopCtl *h_ctl;   // pinned host: [0] the last full copy, [1..2] the flags of
                      // the two batches in flight (f only)
    hipEvent_t bev[2];
    const int *skip;  // &ctl->f.done while iterations are queued ahead, else null
    int loop_batch;   // iterations per queued batch (0: one at a time, waited for)

    // kernel timing (HIP events), see evpgpu_set_timing; one set of pairs per
    // slot, the slot of the iteration being queued in tslot
    int timing;
    unsigned timing_groups; // the timers recorded (bit k: timer k) ...
    int timing_every;       // ... in one Green iteration of every timing_every
    long long tcount;       // Green iterations queued since timing was set
    int tskip;              // the iteration being queued is not timed
    int tslot;
    unsigned tmask[EVP_NSLOT]; // pairs recorded since the slot's last collect
    hipEvent_t ev[EVP_NSLOT][EVP_NTIMERS][2];
    double t_ms[EVP_NTIMERS];
    long long t_n[EVP_NTIMERS];

    // macro state (host)
    evpgpu_macro m;
    double dD[9], dD_acum[9], dD_applied[9];
    double udot[9], udot_s[9], udot_a[9], scauchy[9];
    int creep_flag;
    double disgrad_avg_actual[9];

    // reverse pass (adjoint.hip), allocated by evpgpu_adjoint_record
    struct evpgpu_adj *adj;

    // dislocation-density law (disl.hip), evpgpu_set_dd_law
    int dd_on;
    int dd_first;            // the next Newton launch is the increment's first
    evpgpu_dd_law dd_law;
    evpgpu_dd_phase dd_ph[2];
    double *dd_tab;          // device: per phase chi[144] then cth[8*144]
    double *rho;             // 7 x 12N: SSD, mobile, GND 1-3, forest, parallel
    double *dd_k;            // 4 x 12N: passing stress, 1/(tau_0+tau_cut), gam_o,
                             // first call's passing stress
};

// disl.hip: the dislocation-density law around the Green loop
struct DDParams {
    double c[2][8], shear_G[2], nu[2], q_slip[2], q_bulk[2], bbB[2][EVPGPU_NSYSMX];
    double bp[2][EVPGPU_NSYSMX][3], bn[2][EVPGPU_NSYSMX][3], bb[2][EVPGPU_NSYSMX][3];
    double kT, T_K, L0, D_0, drx, dt;
    int nsys[2], is_gas[2];
    int istep, use_grad;
};
DDParams dd_params(const evpgpu_ctx *c);
int dd_prologue_launch(evpgpu_ctx *c, const DDParams &D);
int dd_evolve_launch(evpgpu_ctx *c, const DDParams &D);

// evpgpu.hip (defined inside the ABI's extern "C" block)
extern "C" int increment_core(evpgpu_ctx *c, int forced, int rep_mix, int rep64, evpgpu_stats *stats);
// adjoint.hip
void adjoint_free(evpgpu_ctx *c);

// ---- launchers (defined in the .hip files) ----
// fft.hip
int fft_forward(evpgpu_ctx *c, const double *real_in, int ncomp, int precision);
int fft_inverse_update(evpgpu_ctx *c, int ncomp, int precision, double *target,
                       const double *add_const, double scale);
int fft_inverse_real(evpgpu_ctx *c, const double2 *cin, int ncomp, double *out, double scale);
int fft_twiddles(evpgpu_ctx *c);
int fft_forward_to(evpgpu_ctx *c, const double *real_in, int ncomp, double2 *dst);
int fft_forward_delta(evpgpu_ctx *c, const double *sig, const float *anchor);
int fft_forward_anchor(evpgpu_ctx *c, const float *anchor);
int fft_inverse_delta(evpgpu_ctx *c, double *E, const double *add, const float *cA, double scale);
int fft_inverse_anchor(evpgpu_ctx *c, float *cA, double scale);
int fft_inverse_set(evpgpu_ctx *c, double *target, const double *base, double scale);
// green.hip
int green_apply(evpgpu_ctx *c, int precision, int accumulate);
int green_skew_fill(evpgpu_ctx *c);
int green_full9(evpgpu_ctx *c, const double2 *in6, double2 *out9);
int green_anchor(evpgpu_ctx *c);
// fft_fused.hip
// HIP-event timing (no-ops unless evpgpu_set_timing is on)
void tmark(evpgpu_ctx *c, int k, int end);
void tcollect(evpgpu_ctx *c);
// collect (keep != 0) or drop the pairs recorded in one slot
void tcollect_slot(evpgpu_ctx *c, int slot, int keep);
// the event pair of timer k in the current slot for a launch that records
// its own start and end (hipExtLaunchKernelGGL), or nulls when not timing
void tkernel(evpgpu_ctx *c, int k, hipEvent_t *start, hipEvent_t *stop);

bool fused_ok(const evpgpu_ctx *c);
// the Green operator's parameters from the context's loading
GreenParams green_params(const evpgpu_ctx *c);
int fused_iteration_fp64(evpgpu_ctx *c, const double *sig, double *E, const double *add, double scale);
int fused_iteration_delta(evpgpu_ctx *c, const double *sig, const float *sigA, double *E, const double *add,
                          const float *cA, double scale);
int fused_anchor(evpgpu_ctx *c, const float *sigA, float *cA, double scale);
int green_delta(evpgpu_ctx *c);
int green_project(evpgpu_ctx *c);
// newton.hip
int newton_launch(evpgpu_ctx *c, const NewtonParams &P, const double *sig_in, double *sig, const double *E,
                  const double *epsp, double *edot, const double *tran, const double *crss,
                  double *gamdot, const int *phase, const int *orient, const double *sgr,
                  long long n, double *partials, int *nblk_out, int *status,
                  double *err_vox, int *niter_vox, int fast, hipEvent_t ev_start = nullptr,
                  hipEvent_t ev_stop = nullptr);
// the Newton kernel's parameters of a context (evpgpu.hip)
NewtonParams context_newton_params(const evpgpu_ctx *c);
// column sums of partials[nblk][nval] into out[0..nval); with status, also
// out[nval] = *status (out must then hold nval + 1 values)
int reduce_partials(evpgpu_ctx *c, const double *partials, int nblk, int nval, double *out,
                    const int *status = nullptr);
// the first stage alone (into c->red2, skipped when *skip is set); returns
// the number of stage-1 rows in *g
int reduce_stage1(evpgpu_ctx *c, const double *partials, int nblk, int nval, const int *skip, int *g);
// loopctl.hip: the device-side Green-loop control
int ctl_init_launch(evpgpu_ctx *c, const LoopCtl &init, const CtlConst &K);
int ctl_phase_launch(evpgpu_ctx *c, int limit, int conv, int first);
// the Newton partials of one iteration -> get_smacro, norms, stopping test
int ctl_iter_launch(evpgpu_ctx *c, int nblk);
// fields.hip
int prologue_launch(evpgpu_ctx *c, const double *dE6, int creep, const double *macro_sym6);
int antisym_add_launch(evpgpu_ctx *c, const double *dW3);
int epilogue_launch(evpgpu_ctx *c, const EpilogueParams &P, double *partials, int *nblk);
int fill_zero(evpgpu_ctx *c, double *p, long long n);
int fill_zero_int(evpgpu_ctx *c, int *p, long long n);
int to_f32_launch(evpgpu_ctx *c, const double *src, float *dst, long long n);
int sub_launch(evpgpu_ctx *c, double *a, const double *b, long long n);
int rsub_launch(evpgpu_ctx *c, const double *a, double *b, long long n);
int copy_launch(evpgpu_ctx *c, const double *src, double *dst, long long n);
