/*
 * evp_oracle.h -- CPU restatement of the EVPFFT Green-operator fixed-point
 * loop (power-law / Voce build and the dislocation-density build) of
 * M-THEMES, used ONLY as the checker.
 *
 * TEST INFRASTRUCTURE.  Nothing in the product links or calls this code:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it, and only to check or time against the HIP path.
 *
 * Every routine cites the reference file:line it restates.  All arithmetic
 * is IEEE double in the reference's operation order; the library is built
 * with -ffp-contract=off so that no multiply-add is fused.
 */
#ifndef EVP_ORACLE_H
#define EVP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NSYSMX 12
#define ORC_NPHMAX 2

/* status codes: the reference's PError/exit codes (see src/evp.h:62-63) */
#define ORC_OK 0
#define ORC_ERR_FILE 1          /* input file not found (src/io.c:103-106)   */
#define ORC_ERR_SINGULAR 35     /* LU row max < 1e-5 (src/kinematics.c:306) */
#define ORC_ERR_LINESEARCH 1203 /* slope >= 0 (src/evolution.c:210-214)     */
#define ORC_ERR_MSINDEX 1139    /* microstructure order (src/init.c:263)    */
#define ORC_ERR_ORTHO 7         /* slip n.b > 1e-3 (src/init.c:195-197)     */
#define ORC_ERR_THET 230        /* thet0 < thet1 (src/init.c:131-133)       */
#define ORC_ERR_BC116 116       /* src/init.c:864-866 */
#define ORC_ERR_BC114 114       /* src/init.c:887-890 */
#define ORC_ERR_NSYS 128        /* nsmx > NSYSMX (assert, src/init.c:128)   */

/* ---- tensor kernels (src/kinematics.c) ---- */
void orc_basis(double B[6][3][3]);
void orc_v6_to_t2(const double v[6], double t[3][3]);         /* chg_basis opt 1 */
void orc_t2_to_v6(const double t[3][3], double v[6]);         /* chg_basis opt 2 */
void orc_c2_to_t4(const double c[6][6], double t4[3][3][3][3]); /* opt 3 */
void orc_t4_to_c2(const double t4[3][3][3][3], double c[6][6]); /* opt 4 */
void orc_v5_to_t2(const double v[5], double t[3][3]);         /* chg_basis5 opt 1 */
void orc_t2_to_v5(const double t[3][3], double v[5]);         /* chg_basis5 opt 2 */
void orc_euler_to_mat(double t1, double ph, double t2, double a[3][3]);
void orc_ten4_transform_inv(double c1[3][3][3][3], const double a[3][3],
                            const double c2[3][3][3][3]);
int orc_lu_inv66(double c[6][6]);
double orc_vonmises(const double t[3][3]);
void orc_rotate_schmid(const double tranmat[3][3], int nsys,
                       const double schm_xt[][5], double schm_gr[][5]);
void orc_orientation(double tran[3][3], double rot[3][3]);

/* ---- constitutive (src/constitutive.c) ---- */
void orc_strain_rate(const double stress[6], double edot[6], double dedot[6][6],
                     int nsys, const double sc[][5], const double tau_t[][2],
                     const double nsr[], const double gam0[], double gamdot[]);
void orc_harden_voxel(int nsys, const double gamdot[], double dt,
                      const double hard[][ORC_NSYSMX], const double tau[][3],
                      const double thet[][2], double crss[][2], double *gamacum);

/* ---- dislocation-density law: the build the reference ships
 * (src/evpFlags.h: DD_BASED_FLAG, DD_CU, FORWARD_PLUS_BACKWARD,
 * DD_GND_FIRST_ORDER, PF_DRX), before any DRX event ---- */
#define ORC_LAW_POWER 0
#define ORC_LAW_DD 1
#define ORC_ERR_NAN 9001 /* assert(isnan(..)==0) in StrainRate_Orowan aborts */
/* law of the jobs that later orc_job_setup calls create (default power law) */
void orc_set_law(int law);
typedef struct {
    double c1, c2, c3, G, q_slip, kT; /* q_slip = Q_slip MY_J_IN_EV / kT */
    double drx;                        /* DeltaTau_DRX exp((t_last - TimeTot)/dt/tau_DRX) */
    const double *bbB;                 /* Burgers length per system [nm] */
    const double (*chi)[ORC_NSYSMX];
    const double (*cth)[ORC_NSYSMX][ORC_NSYSMX]; /* cos s,g1,g2,g3, sin s,g1,g2,g3 */
} orc_dd_phase;
int orc_strain_rate_dd(const double stress[6], double edot[6], double dedot[6][6], int nsys,
                       const double sc[][5], const orc_dd_phase *ph, const double rho_s[],
                       const double rho_g1[], const double rho_g2[], const double rho_g3[],
                       double rho_m[], double rho_F[], double rho_P[], double gamdot[]);
void orc_dd_interaction(const double inter[6], double chi[12][12]);
void orc_dd_angles(int nsys, const double bn[][3], const double bb[][3], double cth[8][12][12]);
double orc_dd_shear_modulus(const double ec[3]);
double orc_dd_kT(double T_K);

/* ---- per-voxel stress update (src/evolution.c:2-310) ----
 * sig  : in = lambda (stress at start of the Green iteration), out = stress
 * returns status; *n_iter receives the Newton iteration count.           */
int orc_update_stress_voxel(const double sg66[6][6], const double c066[6][6],
                            double dt, double sig[3][3], const double eps[3][3],
                            const double disgrad[3][3], int nsys,
                            const double sc[][5], const double crss[][2],
                            const double nsr[], const double gam0[],
                            double edot_out[3][3], double gamdot[],
                            double *err_s, double *err_e, int *n_iter);

/* Newton tolerances of orc_update_stress_voxel, process-wide: relative step
 * (reference 1e-10) and line-search TOLX (reference 1e-7).  Validation runs
 * tighten them so that the loop output is smooth in the material parameters. */
void orc_set_newton_tol(double tol, double tolx);

/* ---- FFT: unnormalised 3-D complex DFT, sign -1 forward / +1 backward ---- */
void orc_fft3d(double *data /* interleaved re,im; [nx][ny][nz] */, int nx, int ny,
               int nz, int sign);

/* ---- whole job (src/init.c SetupJob, src/evolution.c Evolution) ---- */
typedef struct orc_job orc_job;

orc_job *orc_job_setup(const char *input_path, int *status);
void orc_job_free(orc_job *job);
/* run n increments; per-increment Green iteration counts go to iters[] (may be NULL);
 * forced_iters (may be NULL) forces the count of each increment (parity protocol). */
int orc_job_run(orc_job *job, int n, int *iters, const int *forced_iters);
int orc_job_nvox(const orc_job *job);
int orc_threads(void);
void orc_set_threads(int n);

/* Line-search record of orc_update_stress_voxel (tests).  Row of ORC_LS_ROW
 * doubles per trial point: [0] Newton iteration, [1] lambda, [2] f(lambda),
 * [3] f(0), [4] slope, [5] alamin, [6] outcome (0 sufficient decrease,
 * 1 lambda < alamin: step rejected, 2 quadratic backtrack, 3 cubic branch
 * with |a| < 1e-4, 4 cubic with disc < 0, 5 cubic with b <= 0, 6 cubic with
 * b > 0), [7] next lambda before the 0.1 floor, [8..13] sigma at the start
 * of the iteration (6-D basis), [14..19] the step (after the stpmax cap),
 * [20] test, [21] |step| before the cap, [22] 1 if capped.            */
#define ORC_LS_ROW 24
void orc_ls_trace(double *buf, int cap_rows);
double orc_ls_next_lambda(double alam, double alam2, double f_new, double f2, double f_old,
                          double slope, int *outcome);
int orc_ls_trace_count(void); /* OpenMP threads the oracle uses */
void orc_job_dims(const orc_job *job, int dims[3]);
/* field access: 0 Sig, 1 Eps, 2 Edot, 3 DisGrad, 4 VelGrad, 5 TranMat (all N^3 x 9),
 * 6 crss (N^3 x 12 x 2), 7 gamdot (N^3 x 12), 8 gamacum (N^3); dislocation-
 * density jobs: 9 rho_s, 10 rho_m, 11 rho_g1, 12 rho_g2, 13 rho_g3, 14 rho_F,
 * 15 rho_P (N^3 x 12 each), 16 gradient of gamdot (N^3 x 12 x 3), 17 rho_dipole,
 * 18 rho_forest (N^3 x 12)                                                         */
double *orc_job_field(orc_job *job, int which);
/* Gradient_ShearRate of the job's current gamdot field into field 16 */
int orc_job_dd_gradient(orc_job *job);
/* DislocationEvolution of the job's current fields (one call) */
int orc_job_dd_evolve(orc_job *job);
/* the dislocation-density job's tables: shear modulus [2], Burgers lengths
 * [2][12], sense vectors b x n [2][12][3], interaction matrix [2][12][12],
 * projection tables [2][8][12][12]; returns the law                        */
int orc_job_dd_tables(const orc_job *job, double *shear_G, double *bbB, double *bp, double *chi,
                      double *cth);
int *orc_job_grain(orc_job *job);
int *orc_job_phase(orc_job *job);
/* Newton iterations of each voxel in the last Green iteration */
int *orc_job_newton_counts(orc_job *job);
/* macro scalars: see orc_job_macro() in evp_oracle.c for the layout */
void orc_job_macro(const orc_job *job, double *out /* 72 doubles */);
double orc_job_newton_iters(const orc_job *job); /* total Newton iterations so far */
/* setup products for tests */
void orc_job_setup_info(const orc_job *job, double c066[36], double s066[36],
                        double c0[81], double s0[81]);
int orc_job_cgr(const orc_job *job, double *cgr /* N^3 x 36 */);
/* Gamma:sigma-hat on the full complex spectrum, as the loop applies it          */
void orc_job_green(const orc_job *job, const double *ksig_r, const double *ksig_i,
                   double *out_r, double *out_i); /* all [N^3][3][3] */

#ifdef __cplusplus
}
#endif
#endif
