// Host-side launch interface between the C ABI (hx_api.cpp) and the templated gfx950 kernels
// (hx_kernels_nd*.hip). One Kernel_set exists per (n_dim, row_size).
#pragma once

#include <hip/hip_runtime.h>

#include "hx_device.hpp"

namespace hx {

struct Launch
{
  hipStream_t stream;
  Dev_mesh m;
  const void* ops;   // device Ops<row_size>
  const void* hops;  // host copy of the same Ops<row_size>
  const void* eops;  // device Eops<row_size, double>
  Ns_params ns;      // transport models (Navier-Stokes)
  double* scalar;    // device scratch: [0] time step reduction
  int precision;     // hx_precision
  Dt_fuse fuse;      // next step's Max_dt fused into the Euler Local (slots == nullptr: not fused)
};

struct Kernel_set
{
  hipError_t (*write_face)(const Launch&, int cls);
  hipError_t (*neighbor)(const Launch&, int cls, bool visc);
  hipError_t (*local)(const Launch&, int cls, bool visc, double coef, int stage, int compute_residual, int use_filter);
  hipError_t (*neighbor_reconcile)(const Launch&, int cls);
  hipError_t (*reconcile_ldg_flux)(const Launch&, int cls, double dt, int compute_residual, int use_filter);
  hipError_t (*max_dt)(const Launch&, int cls, bool visc, int is_local, double max_cfl_c, double max_cfl_d);
  hipError_t (*prolong)(const Launch&, int n_var, int scale, int offset);
  hipError_t (*restrict_)(const Launch&, int n_var, int scale, int offset);
  hipError_t (*state_bcs)(const Launch&);
  hipError_t (*flux_bcs)(const Launch&);
  // artificial-viscosity generator and the admissibility fix (kind: HX_PDE_ADVECTION, HX_PDE_SMOOTH_AV or
  // HX_PDE_FIX_ADMIS; the diff_* entries take the two diffusion-only kinds)
  hipError_t (*av_write_face)(const Launch&, int cls, int kind);
  hipError_t (*av_neighbor)(const Launch&, int cls, int kind);
  hipError_t (*adv_local)(const Launch&, int cls, double coef, int stage, int compute_residual, double advect_length);
  hipError_t (*diff_local)(const Launch&, int cls, int kind, double dt, double diff_time);
  hipError_t (*diff_neighbor_reconcile)(const Launch&, int cls, int kind);
  hipError_t (*diff_reconcile)(const Launch&, int cls, int kind, double dt, double diff_time, double cheby);
  hipError_t (*av_max_dt)(const Launch&, int cls, int smooth, int is_local, double max_cfl_c, double max_cfl_d);
  hipError_t (*av_bcs)(const Launch&, int mode);
  hipError_t (*av_points)(const Launch&, int cls, int what, int av_rs, double mult, double us_max, double* resid);
  // batch of independent cases: global-time-step Max_dt with a minimum per case
  hipError_t (*max_dt_cases)(const Launch&, int cls, bool visc, double max_cfl_c, double max_cfl_d, const int* case_id,
                             unsigned long long* case_min);
  // Solver::is_admissible (flags[0]: some state inadmissible, flags[1]: some value not finite; zero them first) and
  // the artificial-viscosity swap of Solver::fix_admissibility
  hipError_t (*admissible)(const Launch&, int* flags);
  hipError_t (*fix_av)(const Launch&, int mode);
  // Solver::calc_jacobian: element jacobians (vtss [n_cartesian + n_deformed][2^nd] receives each element's vertex
  // time step scale), boundary and connection normals, hanging-vertex matching of vtss (mode 0 read, 1 write)
  hipError_t (*set_jacobian)(const Launch&, const double* vert, const double* adj, double* vtss, const double* node,
                             const double* diff, const double* bound);
  hipError_t (*connection_normals)(const Launch&);
  hipError_t (*hanging_match)(const Launch&, const int* hanging, int n_hanging, double* vtss, double* vals, int mode);
};

const Kernel_set* kernel_set_nd1(int rs);
const Kernel_set* kernel_set_nd2(int rs);
const Kernel_set* kernel_set_nd3(int rs);

}  // namespace hx
