// Device-side building blocks shared by the gfx950 kernels: the device mesh view, per-row-size operator
// constants, Connection_direction algebra and the Navier-Stokes point functions (pde.hpp:41-175).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/hexed_amd.h"

namespace hx {

constexpr __host__ __device__ int ipow(int b, int e)
{
  int r = 1;
  for (int i = 0; i < e; ++i) r *= b;
  return r;
}

constexpr double heat_rat = 1.4;               // pde.hpp:51
constexpr double specific_gas_air = 287.05287;  // constants.hpp:44

// Device copy of the mesh (all pointers are device addresses). Mirrors hx_host_mesh.
struct Dev_mesh
{
  int nd, rs, nq, nfq, nv, face_sz;
  int n_elem[2];
  double* state[2];
  double* rcache[2];
  double* tss[2];
  double* av[2];
  double* forcing[2];
  double* advection[2];
  double* nom_size[2];
  double* vertex_tss[2];
  int* face_slot[2];
  double* metrics;
  int* face_nslot;
  int n_slot;
  double* faces;
  int n_nslot;
  double* normals;
  int n_con[2];
  int* con_slot[2];
  int* con_dir[2];
  int* con_nslot;
  int n_ref;
  int* ref_slot;
  int* ref_stretch;
  int n_bc_con;
  int* bc_con;
  int* bc_serial;
  int n_bc;
  hx_bc* bcs;
};

// Operator tables of one row size, passed by value as a kernel argument.
template <int RS>
struct Ops
{
  double diff[RS][RS];   // diff_mat[i][j]
  double bound[2][RS];   // boundary[s][j]
  double lift[RS][2];    // W^{-1} B^T diag(-1, 1)  (Derivative.hpp:27-28)
  double filt[RS][RS];   // filter[i][j]
  double node[RS];
  double weight[RS];     // quadrature weights
  double orth[RS][RS];   // orthogonal[degree][node]: normalized Legendre polynomials at the nodes
  double prolong[2][RS][RS];
  double restrict_[2][RS][RS];
};

// Element-kernel operator constants (Derivative.hpp:14-61, Basis filter), in the arithmetic type T
template <int RS, class T>
struct Eops
{
  T diff[RS][RS];  // D
  T lift[RS][2];   // L = W^{-1} B^T diag(-1, 1)
  T bound[2][RS];  // B
  T filt[RS][RS];  // modal filter
};

template <int RS, class T>
Eops<RS, T> make_eops(const Ops<RS>& o)
{
  Eops<RS, T> e;
  for (int i = 0; i < RS; ++i) {
    for (int k = 0; k < RS; ++k) {
      e.diff[i][k] = T(o.diff[i][k]);
      e.filt[i][k] = T(o.filt[i][k]);
    }
    e.lift[i][0] = T(o.lift[i][0]);
    e.lift[i][1] = T(o.lift[i][1]);
    e.bound[0][i] = T(o.bound[0][i]);
    e.bound[1][i] = T(o.bound[1][i]);
  }
  return e;
}

// Transport_model::coefficient (Transport_model.hpp:154-157); sqrt_ref_temp precomputed on the host
struct Transport
{
  double const_val, ref_val, ref_temp, sqrt_ref_temp, temp_offset;
  template <class T>
  __device__ T coefficient(T sqrt_temp) const
  {
    T r = sqrt_temp / T(sqrt_ref_temp);
    return T(const_val) + T(ref_val) * (((T(1) * r) * r) * r) * (T(ref_temp) + T(temp_offset)) /
                              (sqrt_temp * sqrt_temp + T(temp_offset));
  }
};

struct Ns_params
{
  Transport visc, cond;
};

// Max_dt of the next time step fused into the last stage's Local (the state it writes is the state the next Max_dt
// reads): each workgroup stores its minimum of 1/scale in block_min[block_offset + blockIdx.x] (a plain store; the
// minimum over the workgroups is taken by a separate small reduction). block_min == nullptr: not fused.
struct Dt_fuse
{
  double* block_min;
  long long block_offset;
  int visc;  // include the diffusive scale (Max_dt of Navier-Stokes)
  double max_cfl_c, max_cfl_d;
  Ns_params ns;
};

// Connection_direction (Kernel_connection.hpp:7-37) from the packed form of hx_host_mesh::con_dir
struct Dir
{
  int i_dim0, i_dim1;
  bool sign0, sign1;
  __device__ explicit Dir(int p) : i_dim0{p & 3}, i_dim1{(p >> 2) & 3}, sign0{bool((p >> 4) & 1)}, sign1{bool((p >> 5) & 1)} {}
  __device__ bool flip_normal(int s) const { return (s ? sign1 : sign0) == bool(s); }
  __device__ bool flip_tangential() const { return (i_dim0 != i_dim1) && (flip_normal(0) == flip_normal(1)); }
  __device__ bool transpose() const { return (i_dim0 == 0 && i_dim1 == 2) || (i_dim0 == 2 && i_dim1 == 0); }
};

// Index into face 1's storage of the value Face_permutation::match_faces moves to position i
// (transpose, then flip; Spatial.hpp:85-129). The same index scatters results back (restore).
template <int ND, int RS>
__device__ inline int face1_index(const Dir& dir, int i)
{
  if constexpr (ND == 3) {
    int c = i / RS, r = i % RS;  // column-major (r, c) of the rs x rs map
    bool tr = dir.transpose();
    if (dir.flip_tangential()) {
      bool colwise = (dir.i_dim0 > 3 - dir.i_dim0 - dir.i_dim1) != tr;
      if (colwise) r = RS - 1 - r;
      else c = RS - 1 - c;
    }
    return tr ? r * RS + c : c * RS + r;
  } else if constexpr (ND == 2) {
    return dir.flip_tangential() ? RS - 1 - i : i;
  } else {
    return i;
  }
}

// pde::Navier_stokes::Pde::Computation::compute_flux_conv for one flux direction (pde.hpp:97-122); the
// divisions by the density are one reciprocal. T is the arithmetic type (double, or float in the FP32 form).
template <int ND, class T>
__device__ inline void flux_conv(const T (&u)[ND + 2], const T (&n)[ND], T (&f)[ND + 2])
{
  constexpr T gm1 = T(heat_rat - 1.);
  const T inv_mass = T(1) / u[ND];
  T kin = 0;
  for (int j = 0; j < ND; ++j) kin += u[j] * u[j];
  kin *= T(.5) * inv_mass;
  const T pressure = gm1 * (u[ND + 1] - kin);
  T mf = 0;
  for (int j = 0; j < ND; ++j) mf += u[j] * n[j];
  f[ND] = mf;
  const T vol_flux = mf * inv_mass;
  f[ND + 1] = (u[ND + 1] + pressure) * vol_flux;
  for (int j = 0; j < ND; ++j) f[j] = u[j] * vol_flux + pressure * n[j];
}

// pressure (pde.hpp:97-104) in FP64
template <int ND>
__device__ inline double pressure_of(const double (&u)[ND + 2])
{
  double kin = 0;
  for (int j = 0; j < ND; ++j) kin += u[j] * u[j];
  kin *= .5 / u[ND];
  return (heat_rat - 1.) * (u[ND + 1] - kin);
}

// compute_char_speed (pde.hpp:161-166): sound speed estimated from total energy
template <int ND>
__device__ inline double char_speed(const double (&u)[ND + 2])
{
  const double inv_mass = 1. / u[ND];
  const double sound = sqrt(heat_rat * (heat_rat - 1) * u[ND + 1] * inv_mass);
  double sq = 0;
  for (int j = 0; j < ND; ++j) sq += u[j] * u[j];
  return sound + sqrt(sq) * inv_mass;
}

// compute_scalars_diff + the physical part of compute_flux_diff (pde.hpp:130-155): f[v][j], the viscous flux of
// variable v in physical direction j, from the physical gradient grad[v][j]. The reference's final product with the
// normal matrix (pde.hpp:156) is applied per reference direction by the caller.
template <int ND, class T>
__device__ inline void flux_visc(const Ns_params& p, const T (&u)[ND + 2], T bulk, T lap, const T (&grad)[ND + 2][ND],
                                 T (&f)[ND + 2][ND])
{
  const T mass = u[ND];
  const T inv_mass = T(1) / mass;
  T kin = 0;
  for (int j = 0; j < ND; ++j) kin += u[j] * u[j];
  kin *= T(.5) * inv_mass;
  const T bulk_av = fabs(bulk), lap_av = fabs(lap);
  constexpr T r_factor = T((heat_rat - 1) / specific_gas_air);
  const T sqrt_temp = sqrt(fmax((u[ND + 1] - kin) * inv_mass, T(0)) * r_factor);
  const T mu = p.visc.coefficient(sqrt_temp);
  const T energy_cond = p.cond.coefficient(sqrt_temp) * r_factor;
  T veloc[ND];
  for (int i = 0; i < ND; ++i) veloc[i] = u[i] * inv_mass;
  T vg[ND][ND];
  for (int i = 0; i < ND; ++i)
    for (int j = 0; j < ND; ++j) vg[i][j] = (grad[i][j] - veloc[i] * grad[ND][j]) * inv_mass;
  T trace = 0;
  for (int i = 0; i < ND; ++i) trace += vg[i][i];
  const T dil = (bulk_av * mass - T(2. / 3.) * mu) * trace;
  for (int v = 0; v < ND + 2; ++v)
    for (int j = 0; j < ND; ++j) f[v][j] = -lap_av * grad[v][j];
  const T e_over_m2 = u[ND + 1] * inv_mass * inv_mass;
  for (int j = 0; j < ND; ++j) {
    T vtvg = 0, vts = 0;
    for (int i = 0; i < ND; ++i) {
      const T stress = mu * (vg[i][j] + vg[j][i]) + (i == j ? dil : T(0));
      f[i][j] -= stress;
      vtvg += veloc[i] * vg[i][j];
      vts += veloc[i] * stress;
    }
    const T ieg = -e_over_m2 * grad[ND][j] + grad[ND + 1][j] * inv_mass - vtvg;
    f[ND + 1][j] -= vts + energy_cond * ieg;
  }
}

// compute_diffusivity (pde.hpp:169-174)
template <int ND>
__device__ inline double diffusivity(const Ns_params& p, const double (&u)[ND + 2], double bulk, double lap)
{
  const double inv_mass = 1. / u[ND];
  double kin = 0;
  for (int j = 0; j < ND; ++j) kin += u[j] * u[j];
  kin *= .5 * inv_mass;
  constexpr double r_factor = (heat_rat - 1) / specific_gas_air;
  const double sqrt_temp = sqrt(fmax((u[ND + 1] - kin) * inv_mass, 0.) * r_factor);
  const double mu = p.visc.coefficient(sqrt_temp);
  const double energy_cond = p.cond.coefficient(sqrt_temp) * r_factor;
  return fabs(lap) + fmax(fabs(bulk) + mu * inv_mass, energy_cond * inv_mass);
}

// Row iteration along dimension `i_dim` (Row_index.hpp:14-64)
template <int ND, int RS>
struct Row
{
  static constexpr __device__ int stride(int i_dim) { return ipow(RS, ND - 1 - i_dim); }
  // node index of qpoint q along i_dim
  static __device__ int node(int q, int i_dim) { return (q / stride(i_dim)) % RS; }
  // first qpoint of the row through q
  static __device__ int base(int q, int i_dim) { return q - node(q, i_dim) * stride(i_dim); }
  // face qpoint (row counter) of the row through q
  static __device__ int fq(int q, int i_dim)
  {
    int s = stride(i_dim);
    return (q / (s * RS)) * s + q % s;
  }
  // first qpoint of the row with face qpoint f
  static __device__ int base_of_fq(int f, int i_dim)
  {
    int s = stride(i_dim);
    return (f / s) * s * RS + f % s;
  }
};

}  // namespace hx
