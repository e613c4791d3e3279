/*
 * evp.h -- drop-in entry points of the M-THEMES EVPFFT solver, backed by the
 * MI355X Green-operator loop (evpgpu.h).
 *
 * The first seven functions replace the reference's public interface
 * (src/evp.h:572-596,615), which a `forge input.in` run calls from main
 * (src/evp.c:7-32):
 *   GetNameList   src/io.c:96-164     namelist parser, same keys and quirks
 *   PrintNameList src/io.c:166-211
 *   OpenFiles     src/io.c:213-230    vm.out, err.out
 *   CloseFiles    src/io.c:232-245
 *   SetupJob      src/init.c:727-1205 (power-law build)
 *   Evolution     src/evolution.c:550-1729 (reconstructed power-law driver)
 *   DestroyJob    src/init.c:1207-1399
 * They keep the reference's calling convention (global state, no
 * arguments).  On an error they print the reference's message and exit with
 * its code, as PError/MPI_Abort does (src/evp.h:62-63), unless
 * evp_set_abort_on_error(0) was called; then they return and evp_status()
 * holds the code.
 *
 * The evp_job_* functions are the same steps on an explicit handle, for
 * callers that hold several jobs or want status codes.
 */
#ifndef EVP_DROPIN_H
#define EVP_DROPIN_H

#include "evpgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- the reference's interface (C linkage) ---- */
int GetNameList(char **argv);
void PrintNameList(void);
void OpenFiles(void);
void CloseFiles(void);
void SetupJob(void);
void Evolution(void);
void DestroyJob(void);

void evp_set_abort_on_error(int on);
int evp_status(void);
/* precision of the Green loop for SetupJob: EVPGPU_FP64 (default) or EVPGPU_MIXED */
void evp_set_precision(int precision);
/* Constitutive law of the jobs SetupJob and evp_job_open set up next:
 * EVP_LAW_POWER (default), the power law with Voce hardening of the
 * reference's input files (src/constitutive.c:5-67, 275-326), or EVP_LAW_DD,
 * the dislocation-density law the reference's build compiles in
 * (src/evpFlags.h:7-24; StrainRate_Orowan, DislocationEvolution, copper
 * elastic constants, src/init.c:821-826).  No counterpart call in the
 * reference, which fixes the law at compile time.                           */
#define EVP_LAW_POWER 0
#define EVP_LAW_DD 1
void evp_set_law(int law);

/* ---- handle API ---- */
typedef struct evp_job evp_job;

/* Parse input_path (namelist, .sx files, microstructure), set the job up on
 * the GPU.  Relative file names are looked up in the working directory, then
 * next to input_path.                                                          */
evp_job *evp_job_open(const char *input_path, int precision, int *status);
void evp_job_close(evp_job *job);
evpgpu_ctx *evp_job_ctx(evp_job *job);
int evp_job_dims(const evp_job *job, int dims[3]);
int evp_job_nsteps(const evp_job *job);
/* run n increments; iters[k] gets the Green iterations of increment k,
 * stats (may be NULL) the per-increment statistics.                           */
int evp_job_run(evp_job *job, int n, int forced_iters, int *iters, evpgpu_stats *stats);
int evp_job_macro(evp_job *job, evpgpu_macro *m);
/* grain, phase: int32[N] as read from the microstructure file */
int evp_job_ids(const evp_job *job, int *grain, int *phase);
/* number of distinct grain IDs of non-gas voxels (gID_list, src/init.c:286-352)
 * and, if ids != NULL, the sorted IDs                                          */
int evp_job_grain_list(const evp_job *job, int *ids);
/* setup products: c066, s066 (36), c0, s0 (81) */
int evp_job_setup_info(const evp_job *job, double *c066, double *s066, double *c0, double *s0);
/* restore the state right after setup (fields, CRSS, orientations, macros) */
int evp_job_reset(evp_job *job);
/* EVP_LAW_POWER or EVP_LAW_DD */
int evp_job_law(const evp_job *job);
/* vm.out row of the last increment (TIME EVM EVMP DVM DVMP SVM SVM_1) */
int evp_job_vm_row(const evp_job *job, double row[7]);
/* Calibration parameters of crystalline phase `phase` (0 or 1), in the order
 * EVPGPU_P_TAU0, TAU1, THETA0, THETA1, NRS (evpgpu.h): set them on every
 * slip system as the .sx reader would (tau0 forward = backward), then
 * restore the post-setup state with the new initial CRSS.  No counterpart in
 * the reference (parameters only come from the .sx file there).            */
int evp_job_set_params(evp_job *job, int phase, const double params[5]);
int evp_job_get_params(const evp_job *job, int phase, double params[5]);

/* ---- on-disk dumps (src/io.c:283-541) ----
 * .iout fields: six files <prefix>_<c>_S<step:%06d>.iout (c = 1..6 for
 * 11 22 33 23 13 12), each N native doubles in voxel order.  f6 is 6 x N SoA.
 * Texture: <prefix>_S<step:%06d>.iout, N 48-byte records {double Bunge
 * angles in degrees[3]; int coord[3] (1-based x, y, z); int grain, phase,
 * buffer (0)}, the layout src/tex2out.c reads.                              */
enum { EVP_FIELD_SIG = 0, EVP_FIELD_EPS, EVP_FIELD_EDOT, EVP_FIELD_ELS, EVP_FIELD_DISGRAD };
int evp_write_iout6(const char *prefix, int step, long long n, const double *f6);
int evp_write_texture(const char *prefix, int step, const int dims[3], const double *tran9,
                      const int *grain, const int *phase);
/* text listing of a texture file, one line per record (src/tex2out.c:56-64) */
int evp_texture_to_text(const char *in_path, const char *out_path);
/* EulerToTransMatrix option 1 (src/kinematics.c:187-217): row-major matrix ->
 * Bunge angles in degrees (with the reference's PI = 3.14159265359)        */
void evp_mat_to_euler(const double a9[9], double angles[3]);
/* dump one field of the job's current state (WriteSigMPI ... WriteDisgradMPI) */
int evp_job_write_field(evp_job *job, int which, const char *prefix, int step);
/* dump the texture of the job's current state (WriteTextureMPI) */
int evp_job_write_texture(evp_job *job, const char *prefix, int step);

#ifdef __cplusplus
}
#endif
#endif
