/* evp_host.h -- internal declarations of the C drop-in (libevp). */
#ifndef EVP_HOST_H
#define EVP_HOST_H

#include <stdio.h>

#include "../../../include/evp.h"

#define EVP_NPHMAX 2
#define EVP_NSYSMX 12

/* status codes of the reference's PError/exit calls (besides evpgpu's) */
#define EVP_ERR_FILE 1          /* input file missing (src/io.c:103-106)           */
#define EVP_ERR_ORTHO 7         /* slip systems not orthogonal (src/init.c:195-197) */
#define EVP_ERR_BC114 114       /* src/init.c:887-890 */
#define EVP_ERR_BC116 116       /* src/init.c:864-866 */
#define EVP_ERR_NSYS 128        /* nsmx > NSYSMX (assert, src/init.c:128)          */
#define EVP_ERR_THET 230        /* thet0 < thet1 (src/init.c:131-133)              */
#define EVP_ERR_MSINDEX 1139    /* microstructure order (src/init.c:263-266)       */
#define EVP_ERR_NPH 747         /* N_phases > NPHMAX (assert, src/init.c:747)      */
#define EVP_ERR_DRX 850         /* the DRX nucleation gate opened (e_vm > 4,
                                   src/evolution.c:850): not built here           */

/* namelist values (src/io.c:25-94, power-law build) */
typedef struct {
    int CellDim[3], N_phases, Type_phases[2], N_steps, Update_Flag, Hard_Flag, Tex_Flag;
    int IterMax, PrintControl[2], VelGrad_BC_Flag[6], Stress_BC_Flag[6], ElastBC;
    double TimeStep, Err, VelGrad_BC[6], Stress_BC[6], ElastConst_PhaseI[4], ElastConst_PhaseII[4];
    char Slip_PhaseI[100], Slip_PhaseII[100], initial_ms[100];
    int status[21];
} evp_namelist;

int evp_parse_namelist(const char *path, evp_namelist *nl);

/* namelist keys of the dislocation-density build (src/io.c:47-93) */
typedef struct {
    double a0[2], L0, GND_ScaleRatio, T__K, Q_slip[2], Q_bulk[2];
    double C[8][2];      /* C_1 .. C_8 */
    double inter[6][2];  /* Selfinter, Coplanar, CrossSlip, GlissileJunction, HirthLock,
                            LomerCottrellLock */
    double rho_SSD_initial, D_0, tau_DRX, DeltaTau_DRX;
} evp_dd_keys;
int evp_parse_dd_keys(const char *path, evp_dd_keys *dd);
/* the law's tables for phase ph (12 systems; bn, bb normalised) */
void evp_dd_phase_setup(const evp_dd_keys *dd, int ph, const evpgpu_phase *phase, double shear_G,
                        double poisson_nu, evpgpu_dd_phase *out);
double evp_dd_shear_modulus(const double ec[3]);
void evp_print_namelist(const evp_namelist *nl);
FILE *evp_open_rel(const char *dir, const char *name);
/* PlasticInit: fills the phase description (src/init.c:52-231) */
int evp_parse_sx(const char *dir, const char *name, evpgpu_phase *ph);

/* tensor helpers (src/kinematics.c) */
void evp_t2_to_v6(const double t[3][3], double v[6]);
void evp_v6_to_t2(const double v[6], double t[3][3]);
void evp_t2_to_v5(const double t[3][3], double v[5]);
void evp_v5_to_t2(const double v[5], double t[3][3]);
void evp_c2_to_t4(const double c[6][6], double t4[3][3][3][3]);
void evp_t4_to_c2(const double t4[3][3][3][3], double c[6][6]);
int evp_lu_inv66(double c[6][6]);
void evp_euler_to_mat(double t1, double ph, double t2, double a[3][3]);
void evp_ten4_transform_inv(double c1[3][3][3][3], const double a[3][3], const double c2[3][3][3][3]);
void evp_cubic_stiffness(const double *ec, double c[3][3][3][3]);

#endif
