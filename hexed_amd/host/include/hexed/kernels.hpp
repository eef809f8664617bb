// Free functions of the hot path (kernels.hpp:11-42 in the hexed tree). On a registered mesh each call is one
// stream-ordered device sequence; Kernel_options' stopwatch trees receive the work units of every kernel.
#ifndef HEXED_AMD_KERNELS_HPP_
#define HEXED_AMD_KERNELS_HPP_

#include <functional>

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

}  // namespace hexed
#endif
