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
  const double* dtp; // device-resident time step the stage kernels multiply their coefficient by, or null
  int fold;          // face-flux folding: the Local computes its faces' fluxes, Neighbor covers only mortar connections
};

struct Kernel_set
{
  hipError_t (*write_face)(const Launch&, int cls, int scale_velocity);
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
  hipError_t (*adv_local)(const Launch&, int cls, double coef, int stage, int compute_residual, double advect_length,
                          int project_av_rs);
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
  // Stab_art_visc's element values into vert [n_cartesian + n_deformed][2^nd], and Solver::set_art_visc_admis's
  // interpolation of vertex values into the Laplacian artificial viscosity
  hipError_t (*stab_indicator)(const Launch&, int cls, double char_speed, double* vert);
  hipError_t (*vertex_interp_lap)(const Launch&, int cls, const double* vert);
  // 3D, face-flux folding: Neighbor on the mortar connections fused with Restrict_refined (k_mortar_restrict)
  hipError_t (*mortar_restrict)(const Launch&);
};

// ---- fused launches over a group of meshes (hx_group) ----------------------------------------------------------
constexpr int fused_max_seg = 16;  // (mesh, class) segments of one plan: 8 meshes x 2 classes

// what a fused kernel reads per mesh (device memory)
struct Fused_member
{
  Dev_mesh m;
  const void* eops;  // Eops<rs, double>
  const void* ops;   // Ops<rs>
  Dt_fuse fuse;      // the Euler Local's fused time-step work (stage 1 of a batch step), or none
};

// tiles of one fused launch (device memory): segment k covers tiles [first[k], first[k + 1]) of member[k], class cls[k],
// row size rs[k]
struct Fused_plan
{
  int n_seg, n_tiles;
  int first[fused_max_seg + 1];
  signed char member[fused_max_seg], cls[fused_max_seg], rs[fused_max_seg];
};

struct Fused_launch
{
  hipStream_t stream;
  const Fused_member* mem;  // device
  const Fused_plan* plan;   // device
  int n_tiles;
  unsigned grid;            // workgroups
  int order;                // 0: tiles strided over the grid, 1: each workgroup one contiguous chunk of `chunk` tiles
  int chunk;
  const void* const* host_ops;  // host Ops<rs> tables indexed by row size (the context's)
};

// kinds of fused launch, for Fused_set::tiles
enum { fused_state_bcs = 0, fused_neighbor = 1, fused_local = 2 };

struct Fused_set
{
  bool (*supports)(int rs);
  hipError_t (*state_bcs)(const Fused_launch&);
  hipError_t (*neighbor)(const Fused_launch&);
  hipError_t (*local_euler)(const Fused_launch&, int precision, bool with_def, double coef, const double* dtp, int stage,
                            int compute_residual, int use_filter);
  int (*tiles)(int kind, const Dev_mesh& m, int cls);  // tiles of one (mesh, class) segment
};

// n_dim 3, row sizes 4, 5 and 6
const Fused_set* fused_set_nd3();

const Kernel_set* kernel_set_nd1(int rs);
const Kernel_set* kernel_set_nd2(int rs);
const Kernel_set* kernel_set_nd3(int rs);

}  // namespace hx
