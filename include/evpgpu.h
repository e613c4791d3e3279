/*
 * evpgpu.h -- C ABI of the MI355X (gfx950) EVPFFT Green-operator loop.
 *
 * This is the device seam under the drop-in entry points of evp.h: one
 * context holds the whole voxel grid resident in HBM and runs the loading
 * increments of src/evolution.c:1318-1547 (prologue, Green fixed-point loop,
 * epilogue) with hand-written HIP kernels.  It replaces, on one GPU, the
 * FFTW-MPI slab loop of src/evolution.c:1343-1463 and the per-voxel work of
 * src/evolution.c:44-310, src/kinematics.c:430-463, src/constitutive.c:196-326.
 *
 * Conventions
 *   - Plain C types only; every call returns an int status (EVPGPU_OK = 0, or
 *     the reference's abort code where one exists: 35 singular LU,
 *     1203 line-search round-off), instead of calling MPI_Abort.
 *   - Host arrays are caller-owned, FP64 (grain/phase int32), structure of
 *     arrays: component-major, then voxel index v = (x*ny + y)*nz + z
 *     (the reference's pIDX, src/evp.h:53).
 *       symmetric 2nd-order fields (6 comps): 11 22 33 23 13 12
 *         (the component order of the .iout writers, src/io.c:400-437)
 *       full 3x3 fields (9 comps): row-major a*3+b
 *       crss: (s*2 + d), d = 0 forward / 1 backward CRSS of system s
 *   - Device memory is library-owned.  One context per host thread; all work
 *     is ordered on the context's stream.  No exceptions cross the ABI.
 */
#ifndef EVPGPU_H
#define EVPGPU_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    EVPGPU_OK = 0,
    EVPGPU_ERR_SINGULAR = 35,     /* LU row max < 1e-5, src/kinematics.c:306-313 */
    EVPGPU_ERR_LINESEARCH = 1203, /* slope >= 0, src/evolution.c:210-214          */
    EVPGPU_ERR_ARG = -1,          /* bad argument                                 */
    EVPGPU_ERR_HIP = -2,          /* HIP runtime failure                          */
    EVPGPU_ERR_SIZE = -3,         /* grid size not supported (each length 2..512)  */
    EVPGPU_ERR_NOMEM = -4,        /* device allocation failed                     */
    EVPGPU_ERR_STATE = -5,        /* call out of order                            */
    EVPGPU_ERR_NODEV = -6,        /* no HIP device                                */
    EVPGPU_ERR_NONFINITE = -7,    /* non-finite stress/norm (the reference's
                                     assert(!isnan), src/constitutive.c:617,653,700) */
    EVPGPU_ERR_PF_SUM = 1385,     /* a voxel's order parameters sum below 0.5: the
                                     assert of GetStatistics, src/pf_xrx.c:1383-1386 */
    EVPGPU_ERR_PF_CAP = -8,       /* an order-parameter list, or the grains of a
                                     3x3x3 neighbourhood, beyond EVPGPU_PF_CAP */
    EVPGPU_ERR_PF_GRAIN = -9      /* a grain index outside rho_avg[0 .. n-1]     */
};

enum { EVPGPU_FP64 = 0, EVPGPU_MIXED = 1 };

#define EVPGPU_NSYSMX 12

typedef struct evpgpu_ctx evpgpu_ctx;

/* One phase as PlasticInit leaves it (src/init.c:52-231). */
typedef struct {
    int is_gas;                    /* Type_phases[ph] != 0 */
    int nsys;                      /* nSYS[ph] */
    double nrs[EVPGPU_NSYSMX];     /* rate exponent n per system (nSRS) */
    double gam0[EVPGPU_NSYSMX];    /* reference shear rate */
    double tau[EVPGPU_NSYSMX][3];  /* tau0 forward, tau0 backward, tau1 */
    double thet[EVPGPU_NSYSMX][2]; /* theta0, theta1 */
    double hard[EVPGPU_NSYSMX][EVPGPU_NSYSMX]; /* latent/self hardening */
    double schm_xt[EVPGPU_NSYSMX][5];          /* Schmid tensors, crystal axes, 5-D basis */
    double bn[EVPGPU_NSYSMX][3], bb[EVPGPU_NSYSMX][3]; /* unit normals, directions */
} evpgpu_phase;

/* Loading and macro constants (src/init.c:363-378, 864-900). */
typedef struct {
    double dt;                     /* TimeStep */
    double err_tol;                /* Err */
    int iter_max;                  /* IterMax */
    int update_flag, hard_flag, elast_bc;
    int velgrad_bc_flag[6], stress_bc_flag[6];
    double velgrad_bc[6];          /* already multiplied by the flags */
    double stress_bc[6];
    double c066[36], s066[36];     /* <C_gr> and its inverse, 6x6 B-basis */
    double c0[81], s0[81];         /* the same as 4th-order tensors */
    double wgt_ph1;                /* volume fraction of phase 1 */
    double d_vm0;                  /* von Mises of dev(sym Udot) */
} evpgpu_loading;

/* Macroscopic state carried between increments (src/evolution.c:550-1729). */
typedef struct {
    int istep;                     /* index of the next increment */
    double time_tot, e_vm, evmp, d_vm, dvmp, s_vm, s_vm1, err_e, err_s;
    double sig_avg[9], sig_avg1[9];
    double disgrad_avg[9];         /* DisGradAvg: elastic guess of this increment */
    double disgrad_avg_t[9];       /* DisGradAvg_t (= DisGradAvg_actual at t)     */
    double eps_avg[9], edot_avg[9];
} evpgpu_macro;

typedef struct {
    int iters;                     /* Green iterations of the increment (all precisions) */
    int fp64_iters;                /* of which FP64 correction sweeps (mixed mode) */
    double newton_iters;           /* sum over voxels and iterations */
    double ls_trials;              /* line-search trial evaluations (NR_quality calls) */
    double seconds;                /* Green-loop wall time */
    int stopped;                   /* 1 if the creep stop (e_vm > 0.3) ended the run */
    double err_s, err_e;           /* Err_s, Err_e of the increment's last Green iteration
                                      (src/evolution.c:1449-1455); the loop ended on IterMax
                                      when max(err_s, err_e) > Err */
} evpgpu_stats;

/* Dislocation-density law of one crystal phase: the constitutive law of the
 * build the reference ships (src/evpFlags.h: DD_BASED_FLAG, DD_CU,
 * FORWARD_PLUS_BACKWARD, DD_GND_FIRST_ORDER), with the tables SetupJob
 * forms (src/init.c:179-203, 593-702, 1040-1125).  Needs 12 slip systems. */
typedef struct {
    double c[8];                    /* C_1 .. C_8 */
    double q_slip, q_bulk;          /* activation energies, 1E-19 J (namelist units) */
    double shear_G;                 /* MPa (Hill average of phase I, src/init.c:826) */
    double poisson_nu;              /* ElastConst_PhaseI[3] */
    double bbB[EVPGPU_NSYSMX];      /* Burgers length, nm */
    double bp[EVPGPU_NSYSMX][3];    /* sense vectors b x n, crystal axes */
    double chi[EVPGPU_NSYSMX][EVPGPU_NSYSMX];   /* interaction strengths */
    double cth[8][EVPGPU_NSYSMX][EVPGPU_NSYSMX]; /* cos s,g1,g2,g3 then sin s,g1,g2,g3 */
} evpgpu_dd_phase;

/* Scalars of the dislocation-density law (src/io.c:47-93). */
typedef struct {
    double T_K;                     /* temperature, K */
    double L0;                      /* voxel length of the GND rates, mm */
    double D_0;                     /* climb diffusion coefficient */
    double rho_ssd_initial;         /* initial SSD density, 1E12/m^2 */
    double tau_drx, delta_tau_drx;  /* DRX ageing time and extra passing stress */
} evpgpu_dd_law;

const char *evpgpu_strerror(int status);

/* Create a context for an nx*ny*nz grid, each length in [2, 512].  Powers of
 * two in [8, 256] take the Stockham transforms; other lengths a direct DFT
 * along that axis (O(L) work per element, correct for any length). */
int evpgpu_create(evpgpu_ctx **ctx, int nx, int ny, int nz, int precision);
int evpgpu_destroy(evpgpu_ctx *ctx);
int evpgpu_set_precision(evpgpu_ctx *ctx, int precision);
/* Newton tolerances of the per-voxel stress update: relative step (default
 * 1e-10, src/evolution.c:105-110) and line-search TOLX (default 1e-7,
 * src/evolution.c:166-170).  The defaults reproduce the reference; tighter
 * values make the loop output smooth in the material parameters at the
 * 1e-6 finite-difference scale (TOLX rejects Newton steps below 1e-7
 * relative), for validating derivatives. */
int evpgpu_set_newton_tolerances(evpgpu_ctx *ctx, double tol, double tolx);
/* Newton iterations whose step the reference's line search certainly rejects
 * (TOLX) end after the residual: |step| <= |S0 R| / lambda_min(S0 + S) is
 * below TOLX.  Results are identical either way (the last Newton launch of
 * each increment is repeated in full for the rejected trial's gamdot); off
 * runs every iteration in full.  Default on. */
int evpgpu_set_newton_shortcut(evpgpu_ctx *ctx, int on);
/* Where ny = nz <= 128, a Green iteration's transforms run as three fused
 * passes (z+y plane R2C, x lines with the Green step, y+z plane C2R) instead
 * of seven; the results agree with the seven-pass path to rounding.  Off
 * forces the seven passes.  Default on. */
int evpgpu_set_fused_transforms(evpgpu_ctx *ctx, int on);
/* The Green loop's control (get_smacro, the error norms and the loop test,
 * src/evolution.c:473-547,1354,1415-1458) runs on the device after every
 * iteration.  With the fused transforms the host queues the iterations in
 * batches of `batch` (1..4), two batches in flight, and reads the loop's
 * flags once per batch; the iterations queued past the loop's end do
 * nothing.  0 waits for every iteration.  Results are identical either way.
 * Default 4.  (No counterpart in the reference, whose loop is host code.) */
int evpgpu_set_loop_batch(evpgpu_ctx *ctx, int batch);
/* Mixed precision: iterations between FP64 transforms of the stress anchor
 * (default 16, at least 1).  No counterpart in the reference (FP64 only). */
int evpgpu_set_anchor_interval(evpgpu_ctx *ctx, int iterations);

int evpgpu_set_phases(evpgpu_ctx *ctx, const evpgpu_phase *phases, int nphases);
int evpgpu_set_loading(evpgpu_ctx *ctx, const evpgpu_loading *ld);

/* grain, phase: int32[N]; orient: int32[N] index into sgr (n_orient x 36, the
 * per-orientation LU_inv_66(C_gr), src/evolution.c:71-74), -1 for gas.        */
int evpgpu_upload_microstructure(evpgpu_ctx *ctx, const int *grain, const int *phase,
                                 const int *orient, int n_orient, const double *sgr);

/* Full voxel state.  Any pointer may be NULL on download to skip that field. */
int evpgpu_upload_state(evpgpu_ctx *ctx, const double *sig6, const double *disgrad9,
                        const double *eps6, const double *edot6, const double *tranmat9,
                        const double *crss24, const double *gamacum);
int evpgpu_download_state(evpgpu_ctx *ctx, double *sig6, double *disgrad9, double *eps6,
                          double *edot6, double *tranmat9, double *crss24, double *gamacum,
                          double *gamdot12);

int evpgpu_set_macro(evpgpu_ctx *ctx, const evpgpu_macro *m);
int evpgpu_get_macro(evpgpu_ctx *ctx, evpgpu_macro *m);

/* Switch the context to the dislocation-density law (after
 * evpgpu_set_phases; every crystal phase needs 12 systems) and set the
 * densities to their initial values (src/init.c:1073-1093): rho_s =
 * rho_ssd_initial, the mobile and GND densities 0.  The Newton update then
 * uses StrainRate_Orowan (src/constitutive.c:507-721) in place of
 * StrainRate_eval, and each increment ends with the shear-rate gradient
 * (Hard_Flag, istep > 1; src/constitutive.c:875-1108) and
 * DislocationEvolution (src/constitutive.c:1346-1520) in place of harden. */
int evpgpu_set_dd_law(evpgpu_ctx *ctx, const evpgpu_dd_law *law, const evpgpu_dd_phase *phases,
                      int nphases);
/* Densities of a dislocation-density context, each N x 12 (voxel-major),
 * 1E12/m^2: SSD, mobile, GND types 1-3, forest, parallel.  Any pointer may
 * be NULL. */
int evpgpu_download_dd_state(evpgpu_ctx *ctx, double *rho_s, double *rho_m, double *rho_g1,
                             double *rho_g2, double *rho_g3, double *rho_f, double *rho_p);
/* AvgDislocation (src/constitutive.c:1669-1780): the sums over crystal
 * voxels and systems of rho_F, rho_P, rho_s, rho_m, |rho_g|, rho_dipole and
 * rho_forest, each divided by the voxel count -- the dd.out columns 3-9. */
int evpgpu_dd_averages(evpgpu_ctx *ctx, double avg[7]);

/* One loading increment: src/evolution.c:1318-1547 plus hardening.
 * forced_iters > 0 runs exactly that many Green iterations (parity protocol). */
int evpgpu_increment(evpgpu_ctx *ctx, int forced_iters, evpgpu_stats *stats);

/* Kernel timing with HIP events on the context's stream (off by default).
 * which: 0 per-voxel Newton update, 1 forward FFT, 2 Green step, 3 inverse FFT
 * with strain update, 4 increment prologue/epilogue, 5 anchor refresh of the
 * mixed-precision loop (FP64 transform of sigma_A).  Reports the summed
 * elapsed milliseconds and the launch count since the last reset.            */
int evpgpu_set_timing(evpgpu_ctx *ctx, int on);
/* The timers recorded while timing is on (bit k: timer k), in one Green
 * iteration of every `every` (after evpgpu_set_timing: all timers, every
 * iteration).  Every pair of events costs the stream a few microseconds
 * between kernels, so a throughput measurement records only the timers it
 * reports, on a sample of the iterations. */
int evpgpu_set_timing_groups(evpgpu_ctx *ctx, unsigned groups, int every);
/* Wait for all work of the context's stream and of the device. */
int evpgpu_synchronize(evpgpu_ctx *ctx);
int evpgpu_get_timing(evpgpu_ctx *ctx, int which, double *ms, long long *launches);

/* ---- building blocks on device-resident or host data (tests, tools) ---- */
/* R2C 3-D DFT (exponent sign -1) of a real host field -> half spectrum
 * [nx][ny][nz/2+1] complex interleaved; C2R inverse (sign +1, unnormalised). */
int evpgpu_fft_r2c_host(evpgpu_ctx *ctx, const double *in, double *out_cplx);
int evpgpu_fft_c2r_host(evpgpu_ctx *ctx, const double *in_cplx, double *out);
/* Gamma-hat : sigma-hat on a host half spectrum of the 6 stress components
 * (6 x [nx][ny][nz/2+1] complex) -> 9 components of the displacement-gradient
 * increment in the same layout (row-major a*3+b).                           */
int evpgpu_green_host(evpgpu_ctx *ctx, const double *sig_hat6, double *e_hat9);
/* Per-voxel stress update (src/evolution.c:44-310) on a host batch of n voxels
 * of the context's phases: inputs lambda (sig6), strain (6, symmetric part of the
 * displacement gradient), plastic strain eps6, orientation tranmat9 (crystal ->
 * sample, the Schmid tensors follow from it as update_schmid does), crss24,
 * phase; sgr/orient as for the microstructure.  Outputs sig6 (in place), edot6,
 * gamdot12, err2 (err_s, err_e per voxel) and niter.                          */
int evpgpu_newton_host(evpgpu_ctx *ctx, int n, const int *phase, const int *orient,
                       const double *sgr, int n_orient, double *sig6, const double *strain6,
                       const double *eps6, const double *tranmat9, const double *crss24,
                       double *edot6, double *gamdot12, double *err2, int *niter);

/* The per-voxel stress update of evpgpu_newton_host on caller-owned DEVICE
 * buffers (same layouts, n voxels), allocating nothing: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  The kernel's scratch (block
 * partial sums and the status word) is a file-scope device array of the
 * library, one per GPU, held until the process ends; n is at most 2^21
 * (EVPGPU_ERR_SIZE above).  Phase and orientation indices are checked on the
 * device first (EVPGPU_ERR_ARG).  Runs on the context's stream and waits for
 * it.  Calls that overlap on one GPU share the scratch array and race. */
#define EVPGPU_NOALLOC_MAX_VOXELS (1 << 21)
int evpgpu_newton_device_noalloc(evpgpu_ctx *ctx, int n, const int *phase, const int *orient,
                                 const double *sgr, int n_orient, double *sig6,
                                 const double *strain6, const double *eps6,
                                 const double *tranmat9, const double *crss24, double *edot6,
                                 double *gamdot12, double *err2, int *niter);

/* Diagnostic: copy n doubles (n even) device to device reps times with the
 * library's streaming copy kernel (HBM traffic 16 n bytes per launch) and
 * report the achieved GB/s.  No counterpart in the reference; used to
 * measure achievable HBM bandwidth and to scale hardware traffic counters. */
int evpgpu_copy_probe(evpgpu_ctx *ctx, long long n, int reps, double *gbs);

/* ------------------------------------------------------------ adjoint --
 * Reverse-mode derivative of the macroscopic stress curve with respect to
 * the hardening and rate parameters of each phase, for calibration.  No
 * counterpart in the reference (it has no derivatives); it differentiates
 * the converged equations of every increment (implicit differentiation of
 * the Green fixed point, src/evolution.c:1343-1463, and of UpdateStress,
 * src/evolution.c:44-310) and the increment-to-increment updates
 * (Eps, update_orient, harden: src/evolution.c:1469-1547,
 * src/constitutive.c:196-326).  Requirements: crystalline phases only (no
 * gas voxels), ElastBC = 1, no creep stop.
 *
 * Parameters, per phase p (grad[p*EVPGPU_NPARAM + i]):
 *   TAU0   initial CRSS, forward and backward (the .sx tau0xf = tau0xb)
 *   TAU1, THETA0, THETA1   Voce constants of every system
 *   NRS    rate exponent n of every system                                */
enum { EVPGPU_P_TAU0 = 0, EVPGPU_P_TAU1, EVPGPU_P_THETA0, EVPGPU_P_THETA1, EVPGPU_P_NRS, EVPGPU_NPARAM };

typedef struct {
    int cg_iters;         /* conjugate-gradient iterations, all increments */
    int cg_max_iters;     /* largest count of one increment */
    double cg_rel_resid;  /* largest final relative residual */
    double seconds;       /* wall time of the reverse pass, replays included */
    double replay_seconds;/* of which forward replays from checkpoints */
} evpgpu_adjoint_stats;

/* Run n increments from the current state like evpgpu_increment (natural
 * convergence), keeping a device snapshot of the state every
 * checkpoint_every increments and after the last one.  sig_avg (n x 6,
 * 11 22 33 23 13 12) receives the macroscopic stress after each increment,
 * stats (n) the increment statistics; either may be NULL. */
int evpgpu_adjoint_record(evpgpu_ctx *ctx, int n_increments, int checkpoint_every,
                          double *sig_avg, evpgpu_stats *stats);

/* Reverse pass over the recorded increments.  sig_avg_bar (n x 6): dL/d of
 * each recorded macroscopic stress component (the 23 entry stands for the
 * symmetric pair).  grad (2 x EVPGPU_NPARAM) receives dL/d(parameters) of
 * the two phases.  Each increment's adjoint solve is a conjugate-gradient
 * solve to relative residual cg_tol (at most cg_max iterations).  Segments
 * are replayed from the checkpoints; the context is left in the state after
 * the recorded increments. */
int evpgpu_adjoint_run(evpgpu_ctx *ctx, const double *sig_avg_bar, double cg_tol, int cg_max,
                       double *grad, evpgpu_adjoint_stats *stats);

/* ---------------------------------------------------------------------------
 * Sparse phase-field recrystallisation step: PF_XRX (src/pf_xrx.c:1469-1698)
 * on one GPU.  The order parameters of a voxel are a list of at most
 * EVPGPU_PF_CAP (grain, value) entries with the reference's PFdata rules
 * (src/PFgrid.h:54-94); Update_grid's x-major sweep (src/pf_xrx.c:895-1311)
 * is computed data-parallel with the same results bit for bit, Update_gID,
 * boundary_conditions and mesh_extend as the reference runs them on one
 * rank (periodic halos).  Grain indices g are 1-based; rho_avg[g-1] is the
 * reference's per-grain stored-energy array (src/evolution.c:555).
 * Host arrays: voxel v = (x*Ly + y)*Lz + z; lists as n[v], idx/val[v*CAP+k].
 */
#define EVPGPU_PF_CAP 16

typedef struct evpgpu_pf evpgpu_pf;

/* the arguments and globals PF_XRX reads (src/pf_xrx.c:1469-1504) */
typedef struct {
    double E_gb, M_gb;        /* grain-boundary energy and mobility */
    double mu, bb;            /* Shear_G[0], bb_B[0][0] */
    double alpha;             /* Scale_Fdeform */
    double pf_length_scale;   /* voxel length dx */
    double pf_time_step;      /* phase-field time step dt */
    int init_gid;             /* last grain ID of the microstructure (src/init.c:352) */
} evpgpu_pf_params;

typedef struct {
    int substeps;             /* substeps of the last call */
    int activity_passes;      /* passes of the lookup-activity fixpoint, summed */
    int trigger_passes;       /* passes of the stored-energy trigger fixpoint, summed */
    long long updated;        /* voxels on the update branch, summed */
    long long rho_writes;     /* rho_avg writes, summed */
    double ms;                /* device time of the last call's substeps */
    double ms_update;         /* of which the voxel-update kernel (k_update) */
    double ms_avg;            /* and the stored-energy averages (k_avg) */
    int update_launches;      /* launches of each of those two */
} evpgpu_pf_stats;

/* One grid of Lx x Ly x Lz voxels (each >= 3) and an n_rho_avg-entry
 * rho_avg, resident on the current device. */
int evpgpu_pf_create(evpgpu_pf **pf, int Lx, int Ly, int Lz, int n_rho_avg);
int evpgpu_pf_destroy(evpgpu_pf *pf);

/* PF_XRX (src/pf_xrx.h:49, src/pf_xrx.c:1469) without its MPI arguments: each voxel's list
 * starts as (gID, 1.0), the stored energy as rho, then `steps` substeps of
 * Update_grid, Update_gID, mesh_extend and the GetStatistics check.
 * gID_new receives Update_gID's grain field; rho_avg (n_rho_avg) is read and
 * written as the reference's array. */
int evpgpu_pf_xrx(evpgpu_pf *pf, const evpgpu_pf_params *P, const int *gID, const double *rho,
                  int steps, int *gID_new, double *rho_avg);

/* The substeps from a given state: evpgpu_pf_set_lists loads interior lists
 * (n, idx, val), the grain field GID of the stored-energy averages and the
 * stored energy; evpgpu_pf_steps runs `steps` substeps as evpgpu_pf_xrx does;
 * evpgpu_pf_get_lists reads the lists and stored energy back (any pointer
 * may be NULL). */
int evpgpu_pf_set_lists(evpgpu_pf *pf, const int *n, const int *idx, const double *val,
                        const int *GID, const double *defeng);
int evpgpu_pf_steps(evpgpu_pf *pf, const evpgpu_pf_params *P, int steps, int *gID_new,
                    double *rho_avg);
int evpgpu_pf_get_lists(evpgpu_pf *pf, int *n, int *idx, double *val, double *defeng);
int evpgpu_pf_get_stats(evpgpu_pf *pf, evpgpu_pf_stats *stats);

/* NucleationCheckDRX (src/evp.h:649, src/integrated.c:161-494) in the branch the shipped
 * build compiles (GB_BULGING_ONLY and Nucl_Static_Flag 1), on its first call,
 * one GPU: a voxel at a grain boundary (GB_indicator, src/evolution.c:816-828,
 * from the neighbour checks of src/constitutive.c:790-853) nucleates when
 * rho_tot / 12 exceeds its static strength kappa, rho_tot the sum over its
 * nsys systems of rho_s + rho_m + |(rho_g1, rho_g2, rho_g3)|.  Nuclei take the
 * grain IDs n_grains + 1.. in sweep order and each paints the voxels within
 * |r - r_n|^2 <= nucleus_radius (no periodic wrap), later nuclei over
 * earlier ones (src/integrated.c:318-405).  Densities are component-major
 * (system i of voxel v at i*N + v).  In/out: gid_rex (set to 1 where a grain
 * is new), kc (k_c_field: zeta_kc * k_c at the nuclei).  Out: grain_out,
 * nuclei, rho_tot, the nucleus voxels in sweep order (up to max_new) and
 * their count n_new (EVPGPU_ERR_ARG when it exceeds max_new). */
int evpgpu_drx_nucleate(int Lx, int Ly, int Lz, int nsys, const int *grain_in, const double *rho_s,
                        const double *rho_m, const double *rho_g1, const double *rho_g2,
                        const double *rho_g3, const double *kappa, int n_grains, int nucleus_radius,
                        double zeta_kc, double k_c, int *grain_out, int *gid_rex, int *nuclei,
                        double *kc, double *rho_tot, long long *nucleus_vox, int max_new, int *n_new);

/* The static nucleation strengths kappa_drx[v] = gsl_ran_weibull(r, a[v], b)
 * for v = 0..n-1 in order (src/init.c:1049-1053, src/evolution.c:862-866),
 * from GSL's MT19937 seeded with `seed` (the reference seeds it from
 * time(); seed 0 is GSL's default 4357).  Host code: the draws are one
 * sequential stream.  evpgpu_drx_uniform_pos gives the stream's
 * gsl_rng_uniform_pos values themselves. */
int evpgpu_drx_weibull(unsigned long seed, long long n, const double *a, double b, double *out);
int evpgpu_drx_uniform_pos(unsigned long seed, long long n, double *out);

#ifdef __cplusplus
}
#endif
#endif
