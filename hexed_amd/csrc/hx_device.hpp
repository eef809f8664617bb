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
  double* bc_cache;  // [n_bc_con][nv][nfq] boundary-face state cache
  // Face-flux folding (the update path): each element computes its faces' numerical fluxes itself from its own and its
  // neighbour's extrapolated states. partner[cls][e * 2nd + face] = {neighbour slot, info, connection normal slot, 0}
  // with info = packed direction | side << 6 | connection class << 7 | boundary << 8, or neighbour slot -1 where the
  // flux is computed before the Local (faces of hanging-face mortars, coarse hanging faces). The connections that
  // touch a mortar are listed per class in mortar_con. The state blocks of the faces are double-buffered: block 0 of a
  // slot is read from faces (alt_rd 0) or faces_alt [n_slot][nv][nfq] (alt_rd 1) and written to the one alt_wr names;
  // the LDG blocks stay in faces. alt_rd = alt_wr = 0 everywhere else.
  int* partner[2];
  int n_mortar_con[2];
  int* mortar_con[2];
  double* faces_alt;
  int alt_rd, alt_wr;
  // [n_ref][4]: the connection of each mortar of a refined face, connection class << 30 | index (-1: no mortar)
  int* ref_con;
};

// the state (block 0) of face slot `slot` in the buffer `alt` selects
__device__ inline double* state_ptr(const Dev_mesh& m, int slot, int alt)
{
  return alt ? m.faces_alt + size_t(slot) * m.nv * m.nfq : m.faces + size_t(slot) * m.face_sz;
}

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
  // The FP32 matrix-core operands of the row contraction, per lane row rr (lane mod 4): entry t * RS + k holds
  // coefficient k of row 4t + rr of the stacked [D; B] (zero past its rs + 2 rows), entry MF_TA * RS + 2t + s holds
  // L[4t + rr][s] (zero past row rs - 1); each lane row padded to whole 16-byte vectors
  static constexpr int MF_TA = (RS + 5) / 4, MF_TL = (RS + 3) / 4, MF_N = MF_TA * RS + 2 * MF_TL;
  static constexpr int MF_STRIDE = (MF_N + 3) / 4 * 4;
  float mfma[4][MF_STRIDE];
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
  using E = Eops<RS, T>;
  for (int rr = 0; rr < 4; ++rr) {
    for (int c = 0; c < E::MF_STRIDE; ++c) e.mfma[rr][c] = 0.f;
    for (int t = 0; t < E::MF_TA; ++t)
      for (int k = 0; k < RS; ++k) {
        const int row = 4 * t + rr;
        if (row < RS) e.mfma[rr][t * RS + k] = float(o.diff[row][k]);
        else if (row < RS + 2) e.mfma[rr][t * RS + k] = float(o.bound[row - RS][k]);
      }
    for (int t = 0; t < E::MF_TL; ++t)
      for (int s = 0; s < 2; ++s)
        if (4 * t + rr < RS) e.mfma[rr][E::MF_TA * RS + 2 * t + s] = float(o.lift[4 * t + rr][s]);
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
  double* block_min;  // one minimum of 1/scale per workgroup (global time step), or null
  long long block_offset;
  int visc;  // include the diffusive scale (Max_dt of Navier-Stokes)
  double max_cfl_c, max_cfl_d;
  Ns_params ns;
  double* point_scale[2];  // per class, each point's 1/scale (the batch's per-case time steps), or null
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

// the face-0 point whose value face1_index moves to face-1 storage position j (the inverse permutation: the transpose
// and the flip undone in reverse order)
template <int ND, int RS>
__device__ inline int face0_index(const Dir& dir, int j)
{
  if constexpr (ND == 3) {
    const bool tr = dir.transpose();
    int c = tr ? j % RS : j / RS, r = tr ? j / RS : j % RS;
    if (dir.flip_tangential()) {
      bool colwise = (dir.i_dim0 > 3 - dir.i_dim0 - dir.i_dim1) != tr;
      if (colwise) r = RS - 1 - r;
      else c = RS - 1 - c;
    }
    return c * RS + r;
  } else if constexpr (ND == 2) {
    return dir.flip_tangential() ? RS - 1 - j : j;
  } else {
    return j;
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

// ---- characteristic decomposition of the Euler flux (pde.hpp:182-260) ----------------------------------------
// About a reference state and a flux direction: eigenvalues u_n - c, u_n + c, u_n of the flux Jacobian, and the
// 1D eigenvectors (rows: normal momentum, mass, energy; one column per eigenvalue, unit mass perturbation).
// decomp() splits a state into three parts, one per eigenvalue, that sum to it; the 1D coordinates come from a
// 3x3 solve, here Gaussian elimination with complete pivoting that sets an unknown with a zero pivot to zero (the
// reference's column-pivoting QR also returns such a basic solution for a singular system, e.g. zero pressure).
__host__ __device__ inline void solve3(const double (&a_in)[3][3], const double (&b_in)[3], double (&x)[3])
{
  double a[3][3], b[3];
  int col[3] = {0, 1, 2};
  for (int i = 0; i < 3; ++i) {
    b[i] = b_in[i];
    for (int j = 0; j < 3; ++j) a[i][j] = a_in[i][j];
  }
  double amax = 0.;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) amax = fmax(amax, fabs(a[i][j]));
  int rank = 0;
  for (int k = 0; k < 3; ++k) {
    int pi = k, pj = k;
    for (int i = k; i < 3; ++i)
      for (int j = k; j < 3; ++j)
        if (fabs(a[i][j]) > fabs(a[pi][pj])) pi = i, pj = j;
    if (!(fabs(a[pi][pj]) > 1e-14 * amax)) break;
    for (int j = 0; j < 3; ++j) {
      const double t = a[k][j];
      a[k][j] = a[pi][j];
      a[pi][j] = t;
    }
    const double tb = b[k];
    b[k] = b[pi];
    b[pi] = tb;
    for (int i = 0; i < 3; ++i) {
      const double t = a[i][k];
      a[i][k] = a[i][pj];
      a[i][pj] = t;
    }
    const int tc = col[k];
    col[k] = col[pj];
    col[pj] = tc;
    for (int i = k + 1; i < 3; ++i) {
      const double f = a[i][k] / a[k][k];
      for (int j = k; j < 3; ++j) a[i][j] -= f * a[k][j];
      b[i] -= f * b[k];
    }
    ++rank;
  }
  double y[3] = {0., 0., 0.};
  for (int k = rank - 1; k >= 0; --k) {
    double r = b[k];
    for (int j = k + 1; j < rank; ++j) r -= a[k][j] * y[j];
    y[k] = r / a[k][k];
  }
  for (int k = 0; k < 3; ++k) x[col[k]] = y[k];
}

template <int ND>
struct Characteristics
{
  double vals[3];
  double vecs[3][3];
  double dir[ND];
  double veloc[ND];
  __host__ __device__ Characteristics(const double (&state)[ND + 2], const double (&normal)[ND])
  {
    double nn = 0.;
    for (int d = 0; d < ND; ++d) nn += normal[d] * normal[d];
    nn = sqrt(nn);
    for (int d = 0; d < ND; ++d) dir[d] = normal[d] / nn;
    const double mass = state[ND];
    double vsq = 0., vn = 0.;
    for (int d = 0; d < ND; ++d) {
      veloc[d] = state[d] / mass;
      vsq += veloc[d] * veloc[d];
      vn += dir[d] * veloc[d];
    }
    const double pres = .4 * (state[ND + 1] - .5 * mass * vsq);
    const double sound = sqrt(1.4 * fmax(pres, 0.) / mass);
    vals[2] = vn;
    vals[0] = vn - sound;
    vals[1] = vn + sound;
    for (int sign = 0; sign < 2; ++sign) {
      const double d_veloc = (2 * sign - 1) * sound / mass, d_pres = 1.4 * pres / mass;
      vecs[0][sign] = vals[2] + mass * d_veloc;
      vecs[1][sign] = 1.;
      vecs[2][sign] = d_pres / .4 + .5 * vsq + mass * vals[2] * d_veloc;
    }
    vecs[0][2] = vals[2];
    vecs[1][2] = 1.;
    vecs[2][2] = .5 * vsq;
  }
  // d[var][eigenvalue]
  __host__ __device__ void decomp(const double (&state)[ND + 2], double (&d)[ND + 2][3]) const
  {
    double mn = 0., vtn = 0.;
    for (int k = 0; k < ND; ++k) {
      mn += dir[k] * state[k];
      vtn += dir[k] * veloc[k];
    }
    double vt[ND], corr[ND], vcorr = 0.;
    for (int k = 0; k < ND; ++k) {
      vt[k] = veloc[k] - dir[k] * vtn;                           // tangential velocity
      corr[k] = (state[k] - dir[k] * mn) - state[ND] * vt[k];    // tangential momentum not induced by mass
      vcorr += veloc[k] * corr[k];
    }
    const double s1d[3] = {mn, state[ND], state[ND + 1] - vcorr};
    double basis[3];
    solve3(vecs, s1d, basis);
    for (int c = 0; c < 3; ++c) {
      d[ND][c] = vecs[1][c] * basis[c];
      d[ND + 1][c] = vecs[2][c] * basis[c];
      for (int k = 0; k < ND; ++k) d[k][c] = dir[k] * (vecs[0][c] * basis[c]) + vt[k] * basis[c];
    }
    for (int k = 0; k < ND; ++k) d[k][2] += corr[k];
    d[ND + 1][2] += vcorr;
  }
};

// Riemann_invariants' apply_char (Boundary_condition.cpp:83-95): decompose `inside` and `outside` about `state`
// and take the outside part of each characteristic whose eigenvalue has the sign `sign`
template <int ND>
__host__ __device__ inline void apply_char(const double (&state)[ND + 2], const double (&normal)[ND], int sign,
                                           const double (&inside)[ND + 2], const double (&outside)[ND + 2],
                                           double (&out)[ND + 2])
{
  const Characteristics<ND> ch(state, normal);
  double di[ND + 2][3], d_o[ND + 2][3];
  ch.decomp(inside, di);
  ch.decomp(outside, d_o);
  for (int v = 0; v < ND + 2; ++v) {
    double sum = 0.;
    for (int c = 0; c < 3; ++c) sum += (sign * ch.vals[c] > 0) ? d_o[v][c] : di[v][c];
    out[v] = sum;
  }
}

// the Stefan-Boltzmann constant from its definition, as constants.hpp:52 computes it (math::pow multiplies from 1)
__host__ __device__ constexpr double ipow_d(double b, int e)
{
  double r = 1.;
  for (int i = 0; i < e; ++i) r *= b;
  return r;
}
constexpr double stefan_boltzmann = 2 * ipow_d(M_PI, 5) * ipow_d(1.380649e-23, 4) /
                                    (15 * ipow_d(299792458., 2) * ipow_d(6.62607015e-34, 3));

// No_slip's thermal condition (Boundary_condition.hpp:140-186, Boundary_condition.cpp:373-379)
template <int NV>
__host__ __device__ inline double ghost_energy(const hx_bc& bc, const double (&state)[NV])
{
  return bc.thermal == HX_THERMAL_ENERGY ? bc.thermal_value * state[NV - 2] : state[NV - 1];
}
template <int NV>
__host__ __device__ inline double ghost_heat_flux(const hx_bc& bc, const double (&state)[NV], double heat_flux)
{
  if (bc.thermal == HX_THERMAL_ENERGY) return heat_flux;
  if (bc.thermal == HX_THERMAL_HEAT_FLUX) return bc.thermal_value;
  const double temp = state[NV - 1] * .4 / state[NV - 2] / specific_gas_air;
  return bc.emissivity * stefan_boltzmann * ipow_d(temp, 4) + bc.heat_transfer_coef * (temp - bc.temperature);
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
