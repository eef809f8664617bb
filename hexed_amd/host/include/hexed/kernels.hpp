// Free functions of the hot path (kernels.hpp:11-42 in the hexed tree). On a registered mesh each call is one
// stream-ordered device sequence; Kernel_options' stopwatch trees receive the work units of every kernel.
#ifndef HEXED_AMD_KERNELS_HPP_
#define HEXED_AMD_KERNELS_HPP_

#include <functional>
#include <memory>

#include "Face_permutation.hpp"
#include "Kernel_connection.hpp"
#include "Kernel_mesh.hpp"
#include "Stopwatch_tree.hpp"
#include "Transport_model.hpp"

namespace hexed
{

struct Kernel_options
{
  Stopwatch_tree& sw_car;
  Stopwatch_tree& sw_def;
  Stopwatch_tree& sw_pr;
  double dt;
  int i_stage;
  bool compute_residual = false;
  bool use_filter = false;
};

void compute_euler(Kernel_mesh, Kernel_options);
//! `flux_bc` is called where the reference calls Solver::apply_flux_bcs; pass an empty function to apply the
//! boundary conditions registered on the device mesh instead
void compute_navier_stokes(Kernel_mesh, Kernel_options, std::function<void()> flux_bc, Transport_model visc,
                           Transport_model therm_cond);
double max_dt_euler(Kernel_mesh, Kernel_options, double convective_safety, double diffusive_safety, bool local_time);
double max_dt_navier_stokes(Kernel_mesh, Kernel_options, double convective_safety, double diffusive_safety,
                            bool local_time, Transport_model visc, Transport_model therm_cond);
void compute_prolong(Kernel_mesh, bool scale = false, bool offset = false);
void compute_restrict(Kernel_mesh, bool scale = true, bool offset = false);
void compute_write_face(Kernel_mesh);

// the artificial-viscosity generator's PDEs (pde::Advection, pde::Smooth_art_visc) and the admissibility fix's
// (pde::Fix_therm_admis); on a registered mesh with the reference element layout, which holds their fields.
// `flux_bc` as for compute_navier_stokes: empty applies the device's own hooks (Solver::apply_avc_diff_flux_bcs,
// Solver::apply_fta_flux_bcs)
void compute_advection(Kernel_mesh, Kernel_options, double advect_length);
void compute_smooth_av(Kernel_mesh, Kernel_options, std::function<void()> flux_bc, double diff_time,
                       double chebyshev_step);
void compute_fix_therm_admis(Kernel_mesh, Kernel_options, std::function<void()> flux_bc);
double max_dt_advection(Kernel_mesh, Kernel_options, double convective_safety, double diffusive_safety,
                        bool local_time, double advect_length);
double max_dt_smooth_av(Kernel_mesh, Kernel_options, double convective_safety, double diffusive_safety,
                        bool local_time);
double max_dt_fix_therm_admis(Kernel_mesh, Kernel_options, double convective_safety, double diffusive_safety,
                              bool local_time);
void compute_prolong_advection(Kernel_mesh);
void compute_write_face_advection(Kernel_mesh);
void compute_write_face_smooth_av(Kernel_mesh);
//! host reordering of one connection's face-1 data (nd + 2 variables) to match face 0, as Neighbor does on the device
std::unique_ptr<Face_permutation_dynamic> face_permutation(int n_dim, int row_size, Connection_direction, double* data);

}  // namespace hexed
#endif
