/*
 * pf_oracle.h -- CPU restatement of M-THEMES' sparse phase-field
 * recrystallisation step (PF_XRX, src/pf_xrx.c:1469-1698) for one rank,
 * used ONLY as the checker.
 *
 * TEST INFRASTRUCTURE.  Nothing in the product links or calls this code:
 * only tests/ and bench.py's cpu_baseline leg load it.
 *
 * The order-parameter store is the reference's PFdata (src/PFgrid.h:54-94):
 * a list of (grain index, value) per cell whose non-const lookup appends a
 * zero entry for an absent grain, or re-labels the last entry when its
 * value is below 5e-3.  The sweep of Update_grid (src/pf_xrx.c:895-1311)
 * reads its neighbours through that lookup, so the restatement keeps the
 * lists, the x-major sweep and every lookup in the reference's order; all
 * arithmetic is IEEE double in the reference's operation order (built with
 * -ffp-contract=off).  The grid is the reference's padded one: three halo
 * planes on every side, filled as boundary_conditions (src/pf_xrx.c:253-681)
 * and mesh_extend (src/pf_xrx.c:683-893) fill them when one rank holds the
 * whole grid.
 */
#ifndef PF_ORACLE_H
#define PF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_PF_CAP 24          /* entries per cell list in the restatement */

#define ORC_PF_OK 0
#define ORC_PF_ERR_SUM 1385    /* a voxel's order parameters sum below 0.5:
                                  the assert of GetStatistics, src/pf_xrx.c:1383-1386 */
#define ORC_PF_ERR_CAP 2001    /* a cell list outgrew ORC_PF_CAP */
#define ORC_PF_ERR_GRAIN 2002  /* a grain index outside rho_avg[0..n-1] (or < 1) */
#define ORC_PF_ERR_ARG 2003

/* the globals and arguments PF_XRX reads (src/pf_xrx.c:1469-1520) */
typedef struct {
    double E_gb, M_gb;        /* grain-boundary energy and mobility          */
    double mu, bb;            /* Shear_G[0], bb_B[0][0]                      */
    double alpha;             /* Scale_Fdeform                               */
    double pf_length_scale;   /* dx                                          */
    double pf_time_step;      /* dt                                          */
    int init_gid;             /* the last grain ID of the microstructure     */
} orc_pf_params;

/* The derived coefficients of src/pf_xrx.c:1497-1504: M_coeff, kappa, m,
 * E_disl, in that order. */
void orc_pf_coeffs(const orc_pf_params *P, double out[4]);

/* PF_XRX on an Lx x Ly x Lz grid (x-major, z fastest): the lists start as
 * one entry (gID, 1.0) per voxel, then `steps` times Update_grid, Update_gID,
 * mesh_extend and the GetStatistics check.  rho_avg (n_rho_avg entries,
 * rho_avg[g-1] for grain g) is read and written as the reference's array.
 * Outputs: gID_new (the grain of the largest order parameter, Update_gID),
 * and when not NULL the final interior lists (n_out[v], idx_out/val_out
 * [v*ORC_PF_CAP + k]) and the stored-energy buffer's interior (defeng_out). */
int orc_pf_xrx(int Lx, int Ly, int Lz, const orc_pf_params *P, const int *gID, const double *rho,
               int steps, double *rho_avg, int n_rho_avg, int *gID_new, int *n_out, int *idx_out,
               double *val_out, double *defeng_out);

/* The same steps from given interior lists (n_in, idx_in, val_in laid out as
 * the outputs), grain field GID and stored energy: the lists are given halos
 * by boundary_conditions, then `steps` substeps run as above. */
int orc_pf_run_lists(int Lx, int Ly, int Lz, const orc_pf_params *P, const int *n_in,
                     const int *idx_in, const double *val_in, const int *GID, const double *defeng,
                     int steps, double *rho_avg, int n_rho_avg, int *gID_new, int *n_out,
                     int *idx_out, double *val_out, double *defeng_out);

/* PFdata's non-const lookup on one list (n entries, capacity cap): returns
 * the position of grain g after the lookup, or -1 when the list is full. */
int orc_pf_lookup(int *n, int cap, int *idx, double *val, int g);

/* laplacian() of src/pf_xrx.c:6-9 with its parameter order */
double orc_pf_laplacian(const double a[19]);

/* NucleationCheckDRX (src/integrated.c:161-494) in the branch the shipped
 * build compiles (GB_BULGING_ONLY, Nucl_Static_Flag == 1), on its first call,
 * one rank.  Densities are component-major (system i of voxel v at
 * i*N + v).  In: grain_in, kappa (kappa_drx), n_grains (gID_list.size()).
 * In/out: gid_rex, kc (k_c_field).  Out: grain_out (grain_f after the
 * nuclei and their spheres), nuclei, rho_tot, the nucleus voxels in sweep
 * order (at most max_new) and their count.  Returns ORC_PF_OK or
 * ORC_PF_ERR_ARG (more nuclei than max_new). */
int orc_drx_nucleate(int Lx, int Ly, int Lz, int nsys, const int *grain_in, const double *rho_s,
                     const double *rho_m, const double *g1, const double *g2, const double *g3,
                     const double *kappa, int n_grains, int nucleus_radius, double zeta_kc,
                     double k_c, int *grain_out, int *gid_rex, int *nuclei, double *kc,
                     double *rho_tot, long long *nucleus_vox, int max_new, int *n_new);

/* GB_indicator (src/evolution.c:816-828) from the grain checks of
 * src/constitutive.c:790-853 on one rank */
void orc_gb_indicator(int Lx, int Ly, int Lz, const int *grain, int *gb);
#ifdef __cplusplus
}
#endif
#endif
