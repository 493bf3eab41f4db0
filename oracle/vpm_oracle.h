/*
 * vpm_oracle.h -- CPU restatement of the reference VPM path (papis84/vpm_free).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity checker for the HIP library in
 * vpm_free_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  Nothing in the product links or calls it.
 *
 * Layout mirrors the Fortran arrays of the reference: grids are component-fastest
 * (RHS_pm(c,i,j,k) -> f[c + nc*(i + nx*(j + ny*k))], 0-based here), particles are
 * XP(3,NVR), QP(neq+1,NVR) column-major.
 *
 * Pinning status per function is listed in vpm_oracle.c's header comment.
 */
#ifndef VPM_ORACLE_H
#define VPM_ORACLE_H
#include <stdint.h>

#define ORC_MAXB 64

/* pm.inp contents, vpm.f90:132-152 */
typedef struct {
    double dpm[3];
    int32_t interf_iproj, ibctyp, idvpm;
    double epsvol;
    int32_t ncoarse, nbi, nbj, nbk, nremesh, ncell_rem, iyntree, ilevmax;
    int32_t ompthreads, idefine, iynslice, ipmwrite;
    int32_t ipmwstart[10], ipmwsteps[10];
} orc_params;

/* Geometry produced by define_sizes (vpm.f90:477-660) */
typedef struct {
    double xbound[6], dpm[3];
    int32_t nn[3], nn_bl[6];
    int32_t nblocks, nbi, nbj, nbk;
    double xbound_bl[6 * ORC_MAXB];
    int32_t nnbl[3 * ORC_MAXB], nnbl_bl[6 * ORC_MAXB];
    double xbound_coarse[6], dpm_coarse[3];
    int32_t nn_coarse[3], nn_bl_coarse[6];
    int32_t ilevmax;
    int32_t defined; /* 1 once a box has been fixed (idefine=1 semantics) */
} orc_geom;

int orc_read_pm_inp(const char *path, orc_params *p);
double orc_projection_fun(int itype, double x);
void orc_definepm(int itype, double xbound[6], double dpm[3], int ndum, const int32_t nsize[3],
                  int32_t nn[3], int32_t nn_bl[6]);
int orc_define_sizes(const orc_params *p, orc_geom *g, const double *xp, int64_t nvr, int ntime);

/* P2M on a grid (xbound,dpm,nn): rhs(ncomp,nn) = sum_p W * qp(c,p); then project_vol3d. */
int orc_project(const double xbound[6], const double dpm[3], const int32_t nn[3], const int32_t nn_bl[6],
                const double *xp, const double *qp, int64_t nvr, int ncomp, int itype, int idvpm,
                double *rhs);

/* Dirichlet 7-point Poisson on window [s..f] (1-based, inclusive) of a grid with nc components.
 * zero_bc=1 -> solve_eq_0_3d semantics (result into out), else solve_eq_3d (BCs from sol). */
void orc_solve_window(const double xbound[6], const double dpm[3], const int32_t nn[3], int nc, int comp,
                      const int32_t win[6], const double *rhs, double *sol, int zero_bc, double *out);

/* Infinite-domain pmesh (iynbc=1, ibctyp=2, direct sums): pmlib.f90:69-192 */
void orc_pmesh_inf(const double xbound[6], const double dpm[3], const int32_t nn[3], int nc,
                   const double *rhs, double *sol);

/* Normal-derivative sources (pinfdomain.f90:294-547) + centroids; returns nbound */
int64_t orc_normalderiv(const double xbound[6], const double dpm[3], const int32_t nn[3], int nc,
                        const double *sol0, double *src /*nc*nbound*/, double *cxyz /*3*nbound*/,
                        double *ds);
/* Boundary sums (pmbound.f90:80-157,415-515), accumulate into sol faces */
void orc_boundsum(const double xbound[6], const double dpm[3], const int32_t nn[3], int nc,
                  int64_t nbound, const double *src, const double *cxyz, const double *ds, double *sol);

/* yaps3d for BLOCKS=1 (yaps3d.f90:1-453): rhs_bl/sol_bl are block arrays (nc, nnbl) */
int orc_yaps3d(const orc_params *p, const orc_geom *g, int nc, const double *rhs_bl, double *sol_bl);

/* calc_velocity_serial_3d(1) (vpm_gcalc.f90:1-136): vel (3 SoA grids), sol overwritten */
void orc_velocity(const orc_geom *g, int neqpm, const double *rhs, double *sol, double *velx,
                  double *vely, double *velz, int idcalc);

/* back_to_particles_3D itype 1 (vpm_time.f90:30-107) */
void orc_m2p(const orc_geom *g, int neqpm, int iproj, const double *sol, const double *velx,
             const double *vely, const double *velz, const double *xp, const double *qp,
             double *up, double *gp, int64_t nvr);

/* vpm(WhatToDo=2) on one block.  Optional field outputs (may be NULL):
 * rhs_out (neqpm+1,NN), sol_out (neqpm,NN: stretching), vel_out (3 grids NN each). */
int orc_vpm2(const orc_params *p, const orc_geom *g, int neqpm, const double *xp, const double *qp,
             double *up, double *gp, int64_t nvr, double *rhs_out, double *psi_out, double *sol_out,
             double *vel_out);

/* remesh_particles_3d(1) with ncell_rem=1 (vpm_remesh.f90:5-225).
 * Pass xp_out=NULL to only count.  Returns the new particle count (or <0 on error). */
int64_t orc_remesh(const orc_params *p, const orc_geom *g, int neqpm, int mrem, const double *xp,
                   const double *qp, int64_t nvr, double *xp_out, double *qp_out);

/* FP64 diagnostics: circ[3] = -sum QP(1:3), imp[3] = 1/2 sum x x (-QP(1:3)) */
void orc_diagnostics(int neqpm, const double *xp, const double *qp, int64_t nvr, double circ[3],
                     double imp[3]);

/* helpers exposed for tests */
void orc_dst1(int m, const double *in, double *out); /* S_k = sum_j x_j sin(pi j k/(m+1)) */
int orc_num_threads(void);
void orc_timers(double *out, int reset);
void orc_set_boundsum_stride(int s);
/* Boundary sums of orc_pmesh_inf by the tree (build_level_nbound_3d + tree_calc_3d) with this
 * levmax (>= 0), or direct (-1). */
void orc_set_tree(int levmax);
void orc_set_yaps_plant(int mode);
void orc_set_block_weights(const int *wx, const int *wy, const int *wz, int nbi, int nbj, int nbk);
void orc_set_tree_factor(double f);
void orc_boundsum_tree(const double xb[6], const double dpm[3], const int32_t nn[3], int nc, int64_t nbound,
                       const double *src, const double *cxyz, const double *ds, double *sol, int levmax);
/* Panel sources (ibctyp = 3): FSOUR_A4 in the frame of target plane iplane (area <= 0: the
 * plane's cell area), the direct sums calc_boundinf_3d_s and the tree sums
 * calc_boundinf_3d_lev_s; orc_set_ibctyp selects them in orc_pmesh_inf. */
double orc_fsour_a4(const double XO[3], const double RG[3], int iplane, const double dpm[3], double area);
void orc_boundsum_s(const double xb[6], const double dpm[3], const int32_t nn[3], int nc, int64_t nbound,
                    const double *src, const double *cxyz, double *sol);
void orc_boundsum_tree_s(const double xb[6], const double dpm[3], const int32_t nn[3], int nc, int64_t nbound,
                         const double *src, const double *cxyz, const double *ds, double *sol, int levmax);
void orc_set_ibctyp(int ibctyp);
void orc_diffuse(const orc_geom *g, int neqpm, double ni, const double *rhs, double *sol);
int orc_vpm45(const orc_params *p, const orc_geom *g, int neqpm, int mode, double ni, const double *xp,
              const double *qp, double *up, double *gp, int64_t nvr, const double *vel, double *sol_out);
void orc_set_threads(int n);

#endif
