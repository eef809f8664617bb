/*
 * TEST INFRASTRUCTURE ONLY. CPU restatement of hexed's DG hot path, used as the parity checker for the
 * HIP implementation and as the CPU baseline in bench.py. Nothing in the product links or calls it.
 *
 * Each function mirrors the reference item named beside it, operating on the same host mesh layout as the
 * product ABI (include/hexed_amd.h). Functions return 0 on success, nonzero with oracle_last_error() set.
 */
#ifndef HEXED_ORACLE_H_
#define HEXED_ORACLE_H_

#include "../include/hexed_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

const char* oracle_last_error(void);
void oracle_set_threads(int n);
int oracle_get_threads(void);

int oracle_write_face(hx_host_mesh* m, int cls, const hx_pde* pde);                        /* Spatial.hpp:29-71   */
int oracle_neighbor(hx_host_mesh* m, int cls, const hx_pde* pde, int i_stage);             /* Spatial.hpp:601-705 */
int oracle_local(hx_host_mesh* m, int cls, const hx_pde* pde, double dt, int i_stage,      /* Spatial.hpp:293-510 */
                 int compute_residual, int use_filter);
int oracle_neighbor_reconcile(hx_host_mesh* m, int cls, const hx_pde* pde);               /* Spatial.hpp:709-760 */
int oracle_reconcile_ldg_flux(hx_host_mesh* m, int cls, const hx_pde* pde, double dt,     /* Spatial.hpp:513-595 */
                              int i_stage, int compute_residual, int use_filter);
int oracle_max_dt(hx_host_mesh* m, int cls, const hx_pde* pde, int is_local, int use_filter, /* Spatial.hpp:763-829 */
                  double safety_conv, double safety_diff, double* dt_out);
int oracle_prolong_refined(hx_host_mesh* m, const hx_pde* pde, int scale, int offset);    /* Spatial.hpp:138-207 */
int oracle_restrict_refined(hx_host_mesh* m, const hx_pde* pde, int scale, int offset);   /* Spatial.hpp:215-286 */
int oracle_apply_state_bcs(hx_host_mesh* m);                                               /* Solver.cpp:56-67    */
int oracle_apply_flux_bcs(hx_host_mesh* m);                                                /* Solver.cpp:69-81    */
/* Characteristics of the Euler flux (pde.hpp:182-260) about `ref` in direction `normal`: vals[3], and parts[nv][3],
 * the decomposition of `state` into the eigenspaces. */
int oracle_characteristics(int n_dim, const double* ref, const double* normal, const double* state, double* vals,
                           double* parts);
int oracle_compute_euler(hx_host_mesh* m, const hx_kernel_options* o);                    /* kernels_convective.cpp:8-18 */
int oracle_compute_navier_stokes(hx_host_mesh* m, const hx_kernel_options* o,             /* kernels_diffusive.cpp:8-29  */
                                 hx_transport visc, hx_transport cond);
int oracle_max_dt_euler(hx_host_mesh* m, const hx_kernel_options* o, double msc, double msd, /* kernels_max_dt.cpp:8-17 */
                        int local_time, double* dt_out);
int oracle_max_dt_navier_stokes(hx_host_mesh* m, const hx_kernel_options* o, double msc, double msd,
                                int local_time, hx_transport visc, hx_transport cond, double* dt_out);
int oracle_update(hx_host_mesh* m, const hx_update_options* o, hx_update_status* st);     /* Solver.cpp:834-886  */
int oracle_compute_residual(hx_host_mesh* m, const hx_update_options* o);                 /* Solver.cpp:899-913  */

/* artificial-viscosity generator (pde.hpp:265-431, Solver.cpp:83-101,428-582); the per-kernel entry points above
 * also take HX_PDE_ADVECTION and HX_PDE_SMOOTH_AV */
int oracle_apply_advection_bcs(hx_host_mesh* m);                                           /* Solver.cpp:508-514  */
int oracle_apply_avc_diff_bcs(hx_host_mesh* m);                                            /* Solver.cpp:83-91    */
int oracle_apply_avc_diff_flux_bcs(hx_host_mesh* m);                                       /* Solver.cpp:93-101   */
int oracle_compute_advection(hx_host_mesh* m, const hx_kernel_options* o, double advect_length);
int oracle_compute_smooth_av(hx_host_mesh* m, const hx_kernel_options* o, double diff_time, double cheby_step);
int oracle_update_art_visc_smoothness(hx_host_mesh* m, const hx_av_options* o, double* residual);

/* thermodynamic admissibility (Solver.cpp:921-1095, thermo.cpp:6-18, pde.hpp:437-493); the per-kernel entry points
 * above also take HX_PDE_FIX_ADMIS. hx_host_mesh carries no vertex connectivity, so the fix needs
 * oracle_set_vertices first: vid_* [n_elem][2^n_dim] shared vertex of each element vertex; hanging [n_hanging][12]:
 * i_dim, is_positive, stretch[2], then (class, element) of each fine element in the hanging face's order. */
int oracle_set_vertices(const hx_host_mesh* m, int n_vertex, const int* vid_cartesian, const int* vid_deformed,
                        int n_hanging, const int* hanging);
int oracle_calc_jacobian(hx_host_mesh* m, const hx_vertex_data* vertices);              /* Solver.cpp:273-381  */
int oracle_is_admissible(hx_host_mesh* m, int* admissible, int* record);                   /* Solver.cpp:921-959  */
int oracle_fix_coefficient(hx_host_mesh* m, const int* record, double* vertex_out);        /* Solver.cpp:999-1030 */
/* Stab_art_visc (stabilizing_art_visc.cpp:8-61): uncert[n_elem[0] + n_elem[1]], Cartesian first */
int oracle_stab_art_visc(hx_host_mesh* m, double char_speed, double* uncert);
/* Solver::set_art_visc_admis (Solver.cpp:636-659) into the Laplacian artificial viscosity (needs oracle_set_vertices);
 * uncert and vertex_out (may be NULL) receive the element values and the shared vertex values */
int oracle_set_art_visc_admis(hx_host_mesh* m, double char_speed, double* uncert, double* vertex_out);
int oracle_compute_fix_therm_admis(hx_host_mesh* m, const hx_kernel_options* o);          /* kernels_diffusive.cpp:32 */
int oracle_max_dt_fix_therm_admis(hx_host_mesh* m, const hx_kernel_options* o, double msc, double msd,
                                  int local_time, double* dt_out);                         /* kernels_max_dt.cpp:21 */
int oracle_fix_admissibility(hx_host_mesh* m, const hx_update_options* o, double stability_ratio,
                             int* n_iters);                                                /* Solver.cpp:961-1095 */

#ifdef __cplusplus
}
#endif
#endif
