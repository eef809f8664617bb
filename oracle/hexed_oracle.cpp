/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle. It restates the reference's DG hot path loop for loop on the
 * host mesh layout of include/hexed_amd.h so that the HIP implementation can be checked against it, and it is
 * the CPU baseline bench.py times. The product never links or calls this file.
 *
 * Reference files restated (paths relative to the hexed source tree):
 *   include/Spatial.hpp          Write_face, Face_permutation, Prolong/Restrict_refined, Local,
 *                                Reconcile_ldg_flux, Neighbor, Neighbor_reconcile, Max_dt
 *   include/pde.hpp:27-175       Navier_stokes<visc>::Pde and its Computation
 *   include/pde.hpp:265-431      Advection and Smooth_art_visc (the artificial-viscosity generator's PDEs)
 *   include/Derivative.hpp       Derivative<row_size>
 *   include/Row_index.hpp, include/Row_rw.hpp   row iteration and row/face access
 *   include/Transport_model.hpp  Sutherland / constant transport coefficients
 *   include/math.hpp:30-36,207-218, src/math.cpp:104-107   pow, interp, chebyshev_step
 *   src/kernels_convective.cpp, src/kernels_diffusive.cpp, src/kernels_max_dt.cpp   kernel sequences
 *   src/Solver.cpp:56-81,117-133,834-913   BC application, max_dt, update, compute_residual
 *   src/Solver.cpp:83-101,428-582          AV boundary conditions, diffuse_art_visc, update_art_visc_smoothness
 *   src/Boundary_condition.cpp   Freestream, Copy, Nonpenetration, Outflow, Riemann_invariants, Pressure_outflow,
 *                                No_slip with its thermal conditions; include/pde.hpp:182-260 Characteristics
 *   include/pde.hpp:437-493      Fix_therm_admis; src/thermo.cpp:6-18 admissible
 *   src/Solver.cpp:40-54,921-1095   vertex sharing, is_admissible, fix_admissibility; src/Hanging_vertex_matcher.cpp,
 *                                src/math.cpp:19-56 (hypercube_matvec), include/math.hpp:188-199 (stretched_ind)
 *
 * What pins each part (closed forms, invariants, the reference's own tests restated) is in the tests under tests/:
 * test_oracle.py, test_basis.py, test_tree.py, test_av.py, test_batch.py, test_bcs.py, test_admissibility.py.
 */
#include "hexed_oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../hexed_amd/csrc/basis_tables.inc"

namespace {

std::string g_err;

constexpr int ipow(int b, int e)
{
  int r = 1;
  for (int i = 0; i < e; ++i) r *= b;
  return r;
}

// math::pow (math.hpp:30-36): repeated multiplication starting from 1
inline double mpow(double base, int e)
{
  double r = 1;
  for (int i = 0; i < e; ++i) r *= base;
  for (int i = 0; i > e; --i) r /= base;
  return r;
}

// math::chebyshev_step (math.cpp:104-107)
inline double chebyshev_step(int n, int i) { return 1 / (1 - std::cos((n - i - 0.5) * M_PI / n)); }

constexpr double heat_rat = 1.4;               // pde.hpp:51
constexpr double specific_gas_air = 287.05287;  // constants.hpp:44

struct Basis
{
  int rs;
  const double* node;
  const double* weight;
  const double* diff;      // [i][j]
  const double* bound;     // [s][j]
  const double* ortho;     // [k][i]
  const double* filt;      // [i][j]
  const double* prol;      // [h][i][j]
  const double* restr;     // [h][i][j]
  const double* eig;       // {min_eig_convection, min_eig_diffusion}
  explicit Basis(int row_size) : rs{row_size}
  {
    int k = row_size - 2;
    node = hx_basis_tables::nodes[k];
    weight = hx_basis_tables::weights[k];
    diff = hx_basis_tables::diff_mats[k];
    bound = hx_basis_tables::boundarys[k];
    ortho = hx_basis_tables::orthogonals[k];
    filt = hx_basis_tables::filters[k];
    prol = hx_basis_tables::prolongs[k];
    restr = hx_basis_tables::restricts[k];
    eig = hx_basis_tables::eigenvalss[k];
  }
  double max_cfl() const { return -2 * 0.9 / eig[0]; }  // Basis.cpp:6-9, quadratic_safety 0.9
  double step_ratio() const { return .5 / 0.9; }        // Basis.cpp:11-14
  double min_eig_diffusion() const { return eig[1]; }
};

// Connection_direction (Kernel_connection.hpp:7-37)
struct Dir
{
  int i_dim[2];
  bool face_sign[2];
  bool flip_normal(int s) const { return face_sign[s] == bool(s); }
  bool flip_tangential() const { return (i_dim[0] != i_dim[1]) && (flip_normal(0) == flip_normal(1)); }
  bool transpose() const { return ((i_dim[0] == 0) && (i_dim[1] == 2)) || ((i_dim[0] == 2) && (i_dim[1] == 0)); }
};

inline Dir unpack_dir(int p) { return Dir{{p & 3, (p >> 2) & 3}, {bool((p >> 4) & 1), bool((p >> 5) & 1)}}; }

// Transport_model (Transport_model.hpp:14-56)
struct Transport
{
  double const_val, ref_val, ref_temp, sqrt_ref_temp, temp_offset;
  bool is_viscous;
  Transport() : const_val{0}, ref_val{0}, ref_temp{1}, sqrt_ref_temp{1}, temp_offset{1}, is_viscous{false} {}
  explicit Transport(const hx_transport& t)
  : const_val{t.const_val}, ref_val{t.ref_val}, ref_temp{t.ref_temp}, sqrt_ref_temp{std::sqrt(t.ref_temp)},
    temp_offset{t.temp_offset}, is_viscous{bool(t.is_viscous)}
  {}
  double coefficient(double sqrt_temp) const
  {
    return const_val + ref_val * mpow(sqrt_temp / sqrt_ref_temp, 3) * (ref_temp + temp_offset) / (sqrt_temp * sqrt_temp + temp_offset);
  }
};

/*
 * pde::Navier_stokes<visc>::Pde<n_dim, row_size> (pde.hpp:41-175). State vector of a Computation holds the
 * n_dim + 2 conserved variables followed by the bulk and Laplacian artificial viscosity coefficients.
 */
template <int nd, bool visc>
struct Navier_stokes
{
  static constexpr bool has_diffusion = visc;
  static constexpr bool has_convection = true;
  static constexpr int n_update = nd + 2;
  static constexpr int n_extrap = nd + 2;
  Transport dyn_visc, therm_cond;

  template <int ndf>
  struct Computation
  {
    const Navier_stokes& eq;
    explicit Computation(const Navier_stokes& e) : eq{e}
    {
      for (int i = 0; i < nd; ++i)
        for (int j = 0; j < ndf; ++j) normal[i][j] = (i == j);
    }
    double state[nd + 4];
    double update_state[nd + 2];
    // fetch_state (pde.hpp:79-84); the AV coefficients live in their own array in this layout
    void fetch_state(int stride, const double* data, const double* av)
    {
      for (int i = 0; i < nd + 2; ++i) state[i] = data[i * stride];
      state[nd + 2] = av ? av[0] : 0.;
      state[nd + 3] = av ? av[stride] : 0.;
    }
    // fetch_extrap_state (pde.hpp:86-92)
    void fetch_extrap_state(int stride, const double* data)
    {
      for (int i = 0; i < nd + 2; ++i) state[i] = data[i * stride];
      state[nd + 2] = 0.;
      state[nd + 3] = 0.;
      for (int i = 0; i < nd + 2; ++i) update_state[i] = state[i];
    }
    double mass, kin_ener, pressure;
    void compute_scalars_conv()  // pde.hpp:97-104
    {
      mass = state[nd];
      kin_ener = 0;
      for (int i = 0; i < nd; ++i) kin_ener += state[i] * state[i];
      kin_ener *= .5 / mass;
      pressure = (heat_rat - 1.) * (state[nd + 1] - kin_ener);
    }
    double normal[nd][ndf];
    double flux_conv[nd + 2][ndf];
    void compute_flux_conv()  // pde.hpp:108-122
    {
      compute_scalars_conv();
      for (int i = 0; i < ndf; ++i) {
        flux_conv[nd][i] = 0;
        for (int j = 0; j < nd; ++j) flux_conv[nd][i] += state[j] * normal[j][i];
        double vol_flux = flux_conv[nd][i] / mass;
        flux_conv[nd + 1][i] = (state[nd + 1] + pressure) * vol_flux;
        for (int j = 0; j < nd; ++j) flux_conv[j][i] = state[j] * vol_flux + pressure * normal[j][i];
      }
    }
    double bulk_av, laplacian_av, sqrt_temp, dyn_visc_coef, therm_cond_coef, energy_cond;
    void compute_scalars_diff()  // pde.hpp:130-138
    {
      bulk_av = std::abs(state[nd + 2]);
      laplacian_av = std::abs(state[nd + 3]);
      sqrt_temp = std::sqrt(std::max((state[nd + 1] - kin_ener) / mass, 0.) * (heat_rat - 1) / specific_gas_air);
      dyn_visc_coef = eq.dyn_visc.coefficient(sqrt_temp);
      therm_cond_coef = eq.therm_cond.coefficient(sqrt_temp);
      energy_cond = therm_cond_coef * (heat_rat - 1) / specific_gas_air;
    }
    double gradient[nd + 2][nd];
    double flux_diff[nd + 2][nd];
    void compute_flux_diff()  // pde.hpp:142-158 (requires ndf == nd)
    {
      static_assert(ndf == nd || ndf == 1);
      if constexpr (ndf == nd) {
        compute_scalars_diff();
        double veloc[nd];
        for (int i = 0; i < nd; ++i) veloc[i] = state[i] / mass;
        double vg[nd][nd];
        for (int i = 0; i < nd; ++i)
          for (int j = 0; j < nd; ++j) vg[i][j] = (gradient[i][j] - veloc[i] * gradient[nd][j]) / mass;
        double trace = 0;
        for (int i = 0; i < nd; ++i) trace += vg[i][i];
        double stress[nd][nd];
        for (int i = 0; i < nd; ++i)
          for (int j = 0; j < nd; ++j)
            stress[i][j] = dyn_visc_coef * (vg[i][j] + vg[j][i]) + (i == j ? (bulk_av * mass - 2. / 3. * dyn_visc_coef) * trace : 0.);
        double fd[nd + 2][nd];
        for (int v = 0; v < nd + 2; ++v)
          for (int j = 0; j < nd; ++j) fd[v][j] = -laplacian_av * gradient[v][j];
        for (int i = 0; i < nd; ++i)
          for (int j = 0; j < nd; ++j) fd[i][j] -= stress[i][j];
        for (int j = 0; j < nd; ++j) {
          double vtvg = 0, vts = 0;
          for (int i = 0; i < nd; ++i) {
            vtvg += veloc[i] * vg[i][j];
            vts += veloc[i] * stress[i][j];
          }
          double int_ener_grad = -state[nd + 1] / mass / mass * gradient[nd][j] + gradient[nd + 1][j] / mass - vtvg;
          fd[nd + 1][j] -= vts + energy_cond * int_ener_grad;
        }
        for (int v = 0; v < nd + 2; ++v)
          for (int j = 0; j < nd; ++j) {
            double s = 0;
            for (int k = 0; k < nd; ++k) s += fd[v][k] * normal[k][j];
            flux_diff[v][j] = s;
          }
      }
    }
    double char_speed;
    void compute_char_speed()  // pde.hpp:161-166
    {
      const double sound_speed = std::sqrt(heat_rat * (heat_rat - 1) * state[nd + 1] / state[nd]);
      double sq = 0;
      for (int i = 0; i < nd; ++i) sq += state[i] * state[i];
      char_speed = sound_speed + std::sqrt(sq) / state[nd];
    }
    double diffusivity;
    void compute_diffusivity()  // pde.hpp:169-174
    {
      compute_scalars_conv();
      compute_scalars_diff();
      diffusivity = std::abs(laplacian_av) + std::max(std::abs(bulk_av) + dyn_visc_coef / mass, energy_cond / mass);
    }
  };
};

// Derivative<row_size> (Derivative.hpp:14-61) with lift = W^{-1} B^T diag(-1, 1)
template <int rs>
struct Derivative
{
  double diff[rs][rs], bound[2][rs], lift[rs][2];
  explicit Derivative(const Basis& b)
  {
    for (int i = 0; i < rs; ++i) {
      for (int j = 0; j < rs; ++j) diff[i][j] = b.diff[i * rs + j];
      for (int s = 0; s < 2; ++s) bound[s][i] = b.bound[s * rs + i];
      double inv_w = 1. / b.weight[i];
      lift[i][0] = -(inv_w * b.bound[0 * rs + i]);
      lift[i][1] = inv_w * b.bound[1 * rs + i];
    }
  }
  // operator()(q, b) = diff*q + lift*(b - boundary*q)   (Derivative.hpp:51-55)
  template <int nv>
  void apply(const double (&q)[rs][nv], const double (&b)[2][nv], double (&out)[rs][nv]) const
  {
    double jump[2][nv];
    for (int s = 0; s < 2; ++s)
      for (int v = 0; v < nv; ++v) {
        double bq = 0;
        for (int k = 0; k < rs; ++k) bq += bound[s][k] * q[k][v];
        jump[s][v] = b[s][v] - bq;
      }
    for (int i = 0; i < rs; ++i)
      for (int v = 0; v < nv; ++v) {
        double dq = 0;
        for (int k = 0; k < rs; ++k) dq += diff[i][k] * q[k][v];
        out[i][v] = dq + (lift[i][0] * jump[0][v] + lift[i][1] * jump[1][v]);
      }
  }
  // operator()(q) = diff*q   (Derivative.hpp:56-60)
  template <int nv>
  void apply(const double (&q)[rs][nv], double (&out)[rs][nv]) const
  {
    for (int i = 0; i < rs; ++i)
      for (int v = 0; v < nv; ++v) {
        double dq = 0;
        for (int k = 0; k < rs; ++k) dq += diff[i][k] * q[k][v];
        out[i][v] = dq;
      }
  }
  // boundary_term(b) = lift*b   (Derivative.hpp:45-49)
  template <int nv>
  void boundary_term(const double (&b)[2][nv], double (&out)[rs][nv]) const
  {
    for (int i = 0; i < rs; ++i)
      for (int v = 0; v < nv; ++v) out[i][v] = lift[i][0] * b[0][v] + lift[i][1] * b[1][v];
  }
};

// Row_index (Row_index.hpp:14-64): rows along i_dim; i_fq enumerates rows (= face quadrature point)
template <int nd, int rs>
struct Rows
{
  static constexpr int nq = ipow(rs, nd);
  static constexpr int nfq = nq / rs;
  static int stride(int i_dim) { return ipow(rs, nd - 1 - i_dim); }
  static int qpoint(int i_dim, int i_fq, int k)
  {
    int s = stride(i_dim);
    return (i_fq / s) * s * rs + i_fq % s + k * s;
  }
  // Row_rw::read_row (Row_rw.hpp:27-35)
  template <int nv>
  static void read_row(const double* data, int i_dim, int i_fq, double (&r)[rs][nv])
  {
    for (int v = 0; v < nv; ++v)
      for (int k = 0; k < rs; ++k) r[k][v] = data[v * nq + qpoint(i_dim, i_fq, k)];
  }
  // Row_rw::write_row (Row_rw.hpp:38-46): d = coef*d + w
  template <int nv>
  static void write_row(const double (&w)[rs][nv], double* data, int i_dim, int i_fq, double coef)
  {
    for (int v = 0; v < nv; ++v)
      for (int k = 0; k < rs; ++k) {
        double& d = data[v * nq + qpoint(i_dim, i_fq, k)];
        d = coef * d + w[k][v];
      }
  }
  // Row_rw::read_bound (Row_rw.hpp:49-57)
  template <int nv>
  static void read_bound(double* const* faces, int i_dim, int i_fq, double (&b)[2][nv])
  {
    for (int v = 0; v < nv; ++v)
      for (int s = 0; s < 2; ++s) b[s][v] = faces[2 * i_dim + s][v * nfq + i_fq];
  }
  // Row_rw::write_bound (Row_rw.hpp:60-68)
  template <int nv>
  static void write_bound(const double (&b)[2][nv], double* const* faces, int i_dim, int i_fq)
  {
    for (int v = 0; v < nv; ++v)
      for (int s = 0; s < 2; ++s) faces[2 * i_dim + s][v * nfq + i_fq] = b[s][v];
  }
};

// Spatial::Face_permutation (Spatial.hpp:73-131) on nv blocks of nfq values
template <int nd, int rs>
struct Face_permutation
{
  static constexpr int nfq = ipow(rs, nd - 1);
  Dir dir;
  double* tgt;
  int n_var;
  void transpose()
  {
    if (dir.transpose() && nd == 3) {
      for (int v = 0; v < n_var; ++v) {
        double* m = tgt + v * nfq;  // column-major rs x rs map: (r, c) at c*rs + r
        for (int r = 0; r < rs; ++r)
          for (int c = r + 1; c < rs; ++c) std::swap(m[c * rs + r], m[r * rs + c]);
      }
    }
  }
  void flip()
  {
    if (!dir.flip_tangential()) return;
    if constexpr (nd == 3) {
      bool colwise = (dir.i_dim[0] > 3 - dir.i_dim[0] - dir.i_dim[1]) != dir.transpose();
      for (int v = 0; v < n_var; ++v) {
        double* m = tgt + v * nfq;
        if (colwise) {
          for (int c = 0; c < rs; ++c)
            for (int r = 0; r < rs / 2; ++r) std::swap(m[c * rs + r], m[c * rs + rs - 1 - r]);
        } else {
          for (int r = 0; r < rs; ++r)
            for (int c = 0; c < rs / 2; ++c) std::swap(m[c * rs + r], m[(rs - 1 - c) * rs + r]);
        }
      }
    } else if constexpr (nd == 2) {
      for (int v = 0; v < n_var; ++v) {
        double* m = tgt + v * nfq;
        for (int r = 0; r < rs / 2; ++r) std::swap(m[r], m[rs - 1 - r]);
      }
    }
  }
  void match_faces() { transpose(); flip(); }
  void restore() { flip(); transpose(); }
};

struct Mesh_view
{
  hx_host_mesh* m;
  int face_sz() const
  {
    int nfq = ipow(m->row_size, m->n_dim - 1);
    return std::max(2 * (m->n_dim + 2), m->n_dim + m->row_size) * nfq;
  }
  double* slot(int s) const { return m->faces + size_t(s) * face_sz(); }
  double* nslot(int s) const { return m->normals + size_t(s) * m->n_dim * ipow(m->row_size, m->n_dim - 1); }
};

template <int nd, int rs>
struct Ops
{
  static constexpr int nq = ipow(rs, nd);
  static constexpr int nfq = nq / rs;
  static constexpr int nv = nd + 2;
  static constexpr int nvert = 1 << nd;
  using R = Rows<nd, rs>;

  // Spatial::Write_face (Spatial.hpp:41-57) for the Navier-Stokes extrapolation variables
  static void write_face_elem(const Basis& basis, const double* read, double* const* faces)
  {
    for (int i_dim = 0; i_dim < nd; ++i_dim)
      for (int i_fq = 0; i_fq < nfq; ++i_fq) {
        double row[rs][nv];
        R::template read_row<nv>(read, i_dim, i_fq, row);
        double b[2][nv];
        for (int s = 0; s < 2; ++s)
          for (int v = 0; v < nv; ++v) {
            double acc = 0;
            for (int k = 0; k < rs; ++k) acc += basis.bound[s * rs + k] * row[k][v];
            b[s][v] = acc;
          }
        R::template write_bound<nv>(b, faces, i_dim, i_fq);
      }
  }

  static void write_face(hx_host_mesh* m, int cls)
  {
    Basis basis(rs);
    Mesh_view mv{m};
    const int n = m->n_elem[cls];
    #pragma omp parallel for
    for (int e = 0; e < n; ++e) {
      double* faces[6];
      for (int f = 0; f < 2 * nd; ++f) faces[f] = mv.slot(m->face_slot[cls][e * 2 * nd + f]);
      write_face_elem(basis, m->state[cls] + size_t(e) * nv * nq, faces);
    }
  }

  template <bool visc>
  static void local(hx_host_mesh* m, int cls, const Navier_stokes<nd, visc>& eq, double dt, int stage,
                    bool compute_residual, bool use_filter)
  {
    using Pde = Navier_stokes<nd, visc>;
    const bool is_deformed = cls == HX_DEFORMED;
    Basis basis(rs);
    Derivative<rs> derivative(basis);
    const double update_coef = stage ? dt * basis.step_ratio() : dt;
    Mesh_view mv{m};
    // dummy face normal for deformed elements in Cartesian connections (Spatial.hpp:331-339)
    static double cartesian_normal[nd][nd][nfq];
    for (int i = 0; i < nd; ++i)
      for (int j = 0; j < nd; ++j)
        for (int q = 0; q < nfq; ++q) cartesian_normal[i][j][q] = i == j;
    const int n = m->n_elem[cls];
    #pragma omp parallel for
    for (int e = 0; e < n; ++e) {
      double* state = m->state[cls] + size_t(e) * nv * nq;
      const double* av = m->av[cls] ? m->av[cls] + size_t(e) * 2 * nq : nullptr;
      double* faces[6];
      double* visc_faces[6];
      for (int f = 0; f < 2 * nd; ++f) {
        faces[f] = mv.slot(m->face_slot[cls][e * 2 * nd + f]);
        visc_faces[f] = faces[f] + nv * nfq;
      }
      double* tss = m->tss[cls] + size_t(e) * nq;
      double d_pos = m->nom_size[cls][e];
      double time_rate[2][nv][nq] = {};
      const double* nrml = nullptr;
      const double* elem_det = nullptr;
      double* face_nrml[6] = {};
      double visc_storage[nd][nv][nq] = {};
      if (is_deformed) {
        nrml = m->metrics + size_t(e) * (nd * nd + 1) * nq;
        elem_det = nrml + nd * nd * nq;
        for (int f = 0; f < 2 * nd; ++f) {
          int ns = m->face_nslot[e * 2 * nd + f];
          face_nrml[f] = ns >= 0 ? mv.nslot(ns) : cartesian_normal[f / 2][0];
        }
      }
      // gradient times Jacobian determinant (Spatial.hpp:371-402)
      if constexpr (Pde::has_diffusion) {
        for (int q = 0; q < nq; ++q)
          for (int v = 0; v < nv; ++v) time_rate[0][v][q] = state[v * nq + q];
        for (int i_dim = 0; i_dim < nd; ++i_dim)
          for (int i_fq = 0; i_fq < nfq; ++i_fq) {
            double row_r[rs][nv];
            R::template read_row<nv>(time_rate[0][0], i_dim, i_fq, row_r);
            double face_state[2][nv];
            R::template read_bound<nv>(visc_faces, i_dim, i_fq, face_state);
            if (is_deformed) {
              double row_n[rs][nd];
              R::template read_row<nd>(nrml + i_dim * nd * nq, i_dim, i_fq, row_n);
              double face_n[2][nd];
              R::template read_bound<nd>(face_nrml, i_dim, i_fq, face_n);
              for (int j_dim = 0; j_dim < nd; ++j_dim)
                for (int v = 0; v < nv; ++v) {
                  double row[rs][1], bound[2][1], d[rs][1];
                  for (int k = 0; k < rs; ++k) row[k][0] = row_n[k][j_dim] * row_r[k][v];
                  for (int s = 0; s < 2; ++s) bound[s][0] = face_n[s][j_dim] * face_state[s][v];
                  derivative.template apply<1>(row, bound, d);
                  for (int k = 0; k < rs; ++k) d[k][0] = d[k][0] / d_pos;
                  R::template write_row<1>(d, visc_storage[j_dim][v], i_dim, i_fq, 1.);
                }
            } else {
              double d[rs][nv];
              derivative.template apply<nv>(row_r, face_state, d);
              for (int k = 0; k < rs; ++k)
                for (int v = 0; v < nv; ++v) d[k][v] = d[k][v] / d_pos;
              R::template write_row<nv>(d, visc_storage[i_dim][0], i_dim, i_fq, 0);
            }
          }
        std::memset(time_rate, 0, sizeof(time_rate));
      }
      // flux (Spatial.hpp:404-442)
      double flux[nd][nv][nq];
      for (int q = 0; q < nq; ++q) {
        typename Pde::template Computation<nd> comp(eq);
        comp.fetch_state(nq, state + q, av ? av + q : nullptr);
        if (is_deformed)
          for (int i_dim = 0; i_dim < nd; ++i_dim)
            for (int j_dim = 0; j_dim < nd; ++j_dim) comp.normal[j_dim][i_dim] = nrml[(i_dim * nd + j_dim) * nq + q];
        comp.compute_flux_conv();
        for (int i_dim = 0; i_dim < nd; ++i_dim)
          for (int v = 0; v < nv; ++v) flux[i_dim][v][q] = comp.flux_conv[v][i_dim];
        if constexpr (Pde::has_diffusion) {
          for (int i_dim = 0; i_dim < nd; ++i_dim)
            for (int v = 0; v < nv; ++v) comp.gradient[v][i_dim] = visc_storage[i_dim][v][q];
          if (is_deformed)
            for (int v = 0; v < nv; ++v)
              for (int i_dim = 0; i_dim < nd; ++i_dim) comp.gradient[v][i_dim] /= elem_det[q];
          comp.compute_flux_diff();
          for (int i_dim = 0; i_dim < nd; ++i_dim)
            for (int v = 0; v < nv; ++v) visc_storage[i_dim][v][q] = comp.flux_diff[v][i_dim];
        }
      }
      // residual (Spatial.hpp:444-470)
      for (int i_dim = 0; i_dim < nd; ++i_dim)
        for (int i_fq = 0; i_fq < nfq; ++i_fq) {
          double row_f[rs][nv], face_f[2][nv], d[rs][nv];
          R::template read_row<nv>(flux[i_dim][0], i_dim, i_fq, row_f);
          R::template read_bound<nv>(faces, i_dim, i_fq, face_f);
          derivative.template apply<nv>(row_f, face_f, d);
          for (int k = 0; k < rs; ++k)
            for (int v = 0; v < nv; ++v) d[k][v] = -d[k][v];
          R::template write_row<nv>(d, time_rate[0][0], i_dim, i_fq, 1.);
          if constexpr (Pde::has_diffusion) {
            R::template read_row<nv>(visc_storage[i_dim][0], i_dim, i_fq, row_f);
            for (int s = 0; s < 2; ++s)
              for (int v = 0; v < nv; ++v) {
                double acc = 0;
                for (int k = 0; k < rs; ++k) acc += basis.bound[s * rs + k] * row_f[k][v];
                face_f[s][v] = acc;
              }
            for (int v = 0; v < nv; ++v)
              for (int s = 0; s < 2; ++s) visc_faces[i_dim * 2 + s][v * nfq + i_fq] = face_f[s][v];
            derivative.template apply<nv>(row_f, d);
            for (int k = 0; k < rs; ++k)
              for (int v = 0; v < nv; ++v) d[k][v] = -d[k][v];
            R::template write_row<nv>(d, time_rate[1][0], i_dim, i_fq, 1.);
          }
        }
      // modal filter (Spatial.hpp:472-481) over both halves of time_rate
      if (use_filter) filter_rows<2 * nv>(basis, time_rate[0][0]);
      // update (Spatial.hpp:483-503)
      double* ref_state = m->rcache[cls] + size_t(e) * std::max(nv, rs) * nq;
      for (int q = 0; q < nq; ++q) {
        double update[nv] = {};
        double mult = update_coef * tss[q] / d_pos;
        if (is_deformed) mult /= elem_det[q];
        for (int v = 0; v < nv; ++v) {
          double u = time_rate[0][v][q];
          if (stage) {
            u -= ref_state[v * nq + q];
          } else {
            ref_state[v * nq + q] = u;
            if constexpr (Pde::has_diffusion) u += time_rate[1][v][q];
          }
          u *= mult;
          if (compute_residual) ref_state[v * nq + q] = u;
          else update[v] = u;
        }
        for (int v = 0; v < nv; ++v) state[v * nq + q] += update[v];
      }
      if constexpr (!Pde::has_diffusion) write_face_elem(basis, state, faces);
    }
  }

  // row-wise modal filter along every dimension: row = filter*row, written with coefficient 0
  template <int nvar>
  static void filter_rows(const Basis& basis, double* data)
  {
    for (int i_dim = 0; i_dim < nd; ++i_dim)
      for (int i_fq = 0; i_fq < nfq; ++i_fq) {
        double row[rs][nvar], out[rs][nvar];
        R::template read_row<nvar>(data, i_dim, i_fq, row);
        for (int i = 0; i < rs; ++i)
          for (int v = 0; v < nvar; ++v) {
            double acc = 0;
            for (int k = 0; k < rs; ++k) acc += basis.filt[i * rs + k] * row[k][v];
            out[i][v] = acc;
          }
        R::template write_row<nvar>(out, data, i_dim, i_fq, 0.);
      }
  }

  // Spatial::Reconcile_ldg_flux (Spatial.hpp:543-594)
  template <bool visc>
  static void reconcile_ldg_flux(hx_host_mesh* m, int cls, double dt, bool compute_residual, bool use_filter)
  {
    const bool is_deformed = cls == HX_DEFORMED;
    Basis basis(rs);
    Derivative<rs> derivative(basis);
    Mesh_view mv{m};
    const int n = m->n_elem[cls];
    #pragma omp parallel for
    for (int e = 0; e < n; ++e) {
      double* state = m->state[cls] + size_t(e) * nv * nq;
      double* faces[6];
      double* visc_faces[6];
      for (int f = 0; f < 2 * nd; ++f) {
        faces[f] = mv.slot(m->face_slot[cls][e * 2 * nd + f]);
        visc_faces[f] = faces[f] + nv * nfq;
      }
      double* tss = m->tss[cls] + size_t(e) * nq;
      double d_pos = m->nom_size[cls][e];
      double time_rate[nv][nq] = {};
      const double* elem_det = is_deformed ? m->metrics + size_t(e) * (nd * nd + 1) * nq + nd * nd * nq : nullptr;
      for (int i_dim = 0; i_dim < nd; ++i_dim)
        for (int i_fq = 0; i_fq < nfq; ++i_fq) {
          double bound_f[2][nv], d[rs][nv];
          R::template read_bound<nv>(visc_faces, i_dim, i_fq, bound_f);
          derivative.template boundary_term<nv>(bound_f, d);
          for (int k = 0; k < rs; ++k)
            for (int v = 0; v < nv; ++v) d[k][v] = -d[k][v];
          R::template write_row<nv>(d, time_rate[0], i_dim, i_fq, 1.);
        }
      if (use_filter) filter_rows<nv>(basis, time_rate[0]);
      double* to_update = compute_residual ? m->rcache[cls] + size_t(e) * std::max(nv, rs) * nq : state;
      for (int q = 0; q < nq; ++q) {
        double mult = dt * tss[q] / d_pos;
        if (is_deformed) mult /= elem_det[q];
        for (int v = 0; v < nv; ++v) to_update[v * nq + q] += time_rate[v][q] * mult;
      }
      write_face_elem(basis, state, faces);
    }
  }

  // Spatial::Neighbor (Spatial.hpp:613-704)
  template <bool visc>
  static void neighbor(hx_host_mesh* m, int cls, const Navier_stokes<nd, visc>& eq)
  {
    using Pde = Navier_stokes<nd, visc>;
    const bool is_deformed = cls == HX_DEFORMED;
    Mesh_view mv{m};
    const int n = m->n_con[cls];
    #pragma omp parallel for
    for (int i_con = 0; i_con < n; ++i_con) {
      Dir dir = unpack_dir(m->con_dir[cls][i_con]);
      double* side[2] = {mv.slot(m->con_slot[cls][2 * i_con]), mv.slot(m->con_slot[cls][2 * i_con + 1])};
      double face[2 + 2 * Pde::has_diffusion][nv * nfq] = {};
      double face_nrml[nd * nfq];
      int sign[2] = {1, 1};
      for (int s = 0; s < 2; ++s)
        for (int i = 0; i < nv * nfq; ++i) face[s][i] = side[s][i];
      Face_permutation<nd, rs> perm{dir, face[1], nv};
      if (is_deformed) {
        perm.match_faces();
        for (int s = 0; s < 2; ++s) sign[s] = 1 - 2 * dir.flip_normal(s);
        const double* nn = mv.nslot(m->con_nslot[2 * i_con]);
        for (int i = 0; i < nd * nfq; ++i) face_nrml[i] = nn[i];
      }
      for (int q = 0; q < nfq; ++q) {
        if constexpr (Pde::has_diffusion) {
          for (int v = 0; v < nv; ++v) {
            double avg = .5 * (face[0][v * nfq + q] + face[1][v * nfq + q]);
            for (int s = 0; s < 2; ++s) face[2 + s][v * nfq + q] = avg;
          }
        }
        typename Pde::template Computation<1> comp[2] = {typename Pde::template Computation<1>(eq),
                                                        typename Pde::template Computation<1>(eq)};
        if (is_deformed) {
          for (int i = 0; i < nd; ++i) comp[0].normal[i][0] = sign[0] * face_nrml[i * nfq + q];
        } else {
          for (int i = 0; i < nd; ++i) comp[0].normal[i][0] = (i == dir.i_dim[0]);
        }
        for (int i = 0; i < nd; ++i) comp[1].normal[i][0] = comp[0].normal[i][0];
        for (int s = 0; s < 2; ++s) comp[s].fetch_extrap_state(nfq, face[s] + q);
        for (int s = 0; s < 2; ++s) {
          comp[s].compute_flux_conv();
          comp[s].compute_char_speed();
        }
        double nsq = 0;
        for (int i = 0; i < nd; ++i) nsq += comp[0].normal[i][0] * comp[0].normal[i][0];
        double coef = std::max(comp[0].char_speed, comp[1].char_speed) * std::sqrt(nsq);
        for (int v = 0; v < nv; ++v) {
          double f = .5 * (comp[0].flux_conv[v][0] + comp[1].flux_conv[v][0] +
                           coef * (comp[0].update_state[v] - comp[1].update_state[v]));
          for (int s = 0; s < 2; ++s) face[s][v * nfq + q] = sign[s] * f;
        }
      }
      if (is_deformed) {
        perm.restore();
        if constexpr (Pde::has_diffusion) Face_permutation<nd, rs>{dir, face[3], nv}.restore();
      }
      for (int s = 0; s < 2; ++s) {
        for (int i = 0; i < nv * nfq; ++i) side[s][i] = face[s][i];
        if constexpr (Pde::has_diffusion)
          for (int i = 0; i < nv * nfq; ++i) side[s][nv * nfq + i] = face[2 + s][i];
      }
    }
  }

  // Spatial::Neighbor_reconcile (Spatial.hpp:716-759)
  static void neighbor_reconcile(hx_host_mesh* m, int cls)
  {
    const bool is_deformed = cls == HX_DEFORMED;
    Mesh_view mv{m};
    const int n = m->n_con[cls];
    #pragma omp parallel for
    for (int i_con = 0; i_con < n; ++i_con) {
      Dir dir = unpack_dir(m->con_dir[cls][i_con]);
      double* side[2] = {mv.slot(m->con_slot[cls][2 * i_con]) + nv * nfq,
                         mv.slot(m->con_slot[cls][2 * i_con + 1]) + nv * nfq};
      double face[2][nv * nfq];
      int sign[2] = {1, 1};
      for (int s = 0; s < 2; ++s)
        for (int i = 0; i < nv * nfq; ++i) face[s][i] = side[s][i];
      Face_permutation<nd, rs> perm{dir, face[1], nv};
      if (is_deformed) {
        perm.match_faces();
        for (int s = 0; s < 2; ++s) sign[s] = 1 - 2 * dir.flip_normal(s);
      }
      for (int q = 0; q < nfq; ++q)
        for (int v = 0; v < nv; ++v) {
          double avg = 0;
          for (int s = 0; s < 2; ++s) avg += .5 * sign[s] * face[s][v * nfq + q];
          for (int s = 0; s < 2; ++s) {
            double& f = face[s][v * nfq + q];
            f = sign[s] * avg - f;
          }
        }
      if (is_deformed) perm.restore();
      for (int s = 0; s < 2; ++s)
        for (int i = 0; i < nv * nfq; ++i) side[s][i] = face[s][i];
    }
  }

  // math::interp (math.hpp:207-218)
  static double interp(double (&values)[nvert], const double (&coords)[nd])
  {
    int stride = nvert;
    for (int i_dim = 0; i_dim < nd; ++i_dim) {
      stride /= 2;
      for (int i = 0; i < stride; ++i) values[i] += coords[i_dim] * (values[i + stride] - values[i]);
    }
    return values[0];
  }

  // Spatial::Max_dt (Spatial.hpp:784-828)
  template <bool visc>
  static double max_dt(hx_host_mesh* m, int cls, const Navier_stokes<nd, visc>& eq, bool is_local, double safety_conv,
                       double safety_diff)
  {
    using Pde = Navier_stokes<nd, visc>;
    Basis basis(rs);
    const double max_cfl_c = basis.max_cfl() * safety_conv;
    const double max_cfl_d = -2 / basis.min_eig_diffusion() * safety_diff;
    double dt = std::numeric_limits<double>::max();
    const int n = m->n_elem[cls];
    #pragma omp parallel for reduction(min : dt)
    for (int e = 0; e < n; ++e) {
      const double* state = m->state[cls] + size_t(e) * nv * nq;
      const double* av = m->av[cls] ? m->av[cls] + size_t(e) * 2 * nq : nullptr;
      double* tss = m->tss[cls] + size_t(e) * nq;
      for (int q = 0; q < nq; ++q) {
        double vertex_spacing[nvert];
        for (int i = 0; i < nvert; ++i) vertex_spacing[i] = m->vertex_tss[cls][size_t(e) * nvert + i];
        double coords[nd];
        for (int i_dim = 0; i_dim < nd; ++i_dim) coords[i_dim] = basis.node[(q / ipow(rs, nd - 1 - i_dim)) % rs];
        double spacing = interp(vertex_spacing, coords);
        typename Pde::template Computation<nd> comp(eq);
        comp.fetch_state(nq, state + q, av ? av + q : nullptr);
        double scale = 0;
        comp.compute_char_speed();
        scale += comp.char_speed / max_cfl_c / spacing;
        if constexpr (Pde::has_diffusion) {
          comp.compute_diffusivity();
          scale += comp.diffusivity / max_cfl_d / spacing / spacing;
        }
        if (is_local) tss[q] = 1. / scale;
        else {
          tss[q] = 1.;
          dt = std::min(dt, 1. / scale);
        }
      }
    }
    return is_local ? 1. : dt;
  }

  // Spatial::Prolong_refined (Spatial.hpp:153-206)
  static void prolong(hx_host_mesh* m, int n_var, bool scl, bool off)
  {
    Basis basis(rs);
    Mesh_view mv{m};
    constexpr int n_face = ipow(2, nd - 1);
    #pragma omp parallel for
    for (int i_ref = 0; i_ref < m->n_ref; ++i_ref) {
      const int* sl = m->ref_slot + 5 * i_ref;
      const bool str[2] = {bool(m->ref_stretch[2 * i_ref]), bool(m->ref_stretch[2 * i_ref + 1])};
      double* coarse = mv.slot(sl[0]) + off * (nd + 2) * nfq;
      int nf = n_face;
      for (int i_dim = 0; i_dim < nd - 1; ++i_dim) nf /= 1 + str[i_dim];
      for (int i_face = 0; i_face < nf; ++i_face) {
        double* fine = mv.slot(sl[1 + i_face]) + off * (nd + 2) * nfq;
        for (int v = 0; v < n_var; ++v) {
          double* var_face = fine + v * nfq;
          for (int q = 0; q < nfq; ++q) var_face[q] = coarse[v * nfq + q];
          for (int i_dim = 0; i_dim < nd - 1; ++i_dim) {
            if (str[i_dim]) {
              for (int q = 0; q < nfq; ++q) var_face[q] *= 1 + scl;
            } else {
              const int pw = nd - 2 - i_dim;
              const int face_stride = str[nd - 2] ? 1 : ipow(2, pw);
              const int qpoint_stride = ipow(rs, pw);
              const int i_half = (i_face / face_stride) % 2;
              const double* mat = basis.prol + i_half * rs * rs;
              for (int i_outer = 0; i_outer < nfq / (rs * qpoint_stride); ++i_outer)
                for (int i_inner = 0; i_inner < qpoint_stride; ++i_inner) {
                  double row[rs], out[rs];
                  for (int k = 0; k < rs; ++k) row[k] = var_face[(i_outer * rs + k) * qpoint_stride + i_inner];
                  for (int i = 0; i < rs; ++i) {
                    double acc = 0;
                    for (int k = 0; k < rs; ++k) acc += mat[i * rs + k] * row[k];
                    out[i] = acc;
                  }
                  for (int k = 0; k < rs; ++k) var_face[(i_outer * rs + k) * qpoint_stride + i_inner] = out[k];
                }
            }
          }
        }
      }
    }
  }

  // Spatial::Restrict_refined (Spatial.hpp:230-285)
  static void restrict_(hx_host_mesh* m, int n_var, bool scl, bool off)
  {
    Basis basis(rs);
    Mesh_view mv{m};
    constexpr int n_face = ipow(2, nd - 1);
    #pragma omp parallel for
    for (int i_ref = 0; i_ref < m->n_ref; ++i_ref) {
      const int* sl = m->ref_slot + 5 * i_ref;
      const bool str[2] = {bool(m->ref_stretch[2 * i_ref]), bool(m->ref_stretch[2 * i_ref + 1])};
      double* coarse = mv.slot(sl[0]) + off * (nd + 2) * nfq;
      for (int i = 0; i < n_var * nfq; ++i) coarse[i] = 0.;
      int nf = n_face;
      for (int i_dim = 0; i_dim < nd - 1; ++i_dim) nf /= 1 + str[i_dim];
      for (int i_face = 0; i_face < nf; ++i_face) {
        double* fine = mv.slot(sl[1 + i_face]) + off * (nd + 2) * nfq;
        for (int v = 0; v < n_var; ++v) {
          double* var_face = fine + v * nfq;
          for (int i_dim = 0; i_dim < nd - 1; ++i_dim) {
            if (str[i_dim]) {
              for (int q = 0; q < nfq; ++q) var_face[q] /= 1 + scl;
            } else {
              const int pw = nd - 2 - i_dim;
              const int face_stride = str[nd - 2] ? 1 : ipow(2, pw);
              const int qpoint_stride = ipow(rs, pw);
              const int i_half = (i_face / face_stride) % 2;
              const double* mat = basis.restr + i_half * rs * rs;
              for (int i_outer = 0; i_outer < nfq / (rs * qpoint_stride); ++i_outer)
                for (int i_inner = 0; i_inner < qpoint_stride; ++i_inner) {
                  double row[rs], out[rs];
                  for (int k = 0; k < rs; ++k) row[k] = var_face[(i_outer * rs + k) * qpoint_stride + i_inner];
                  for (int i = 0; i < rs; ++i) {
                    double acc = 0;
                    for (int k = 0; k < rs; ++k) acc += mat[i * rs + k] * row[k];
                    out[i] = acc;
                  }
                  for (int k = 0; k < rs; ++k) var_face[(i_outer * rs + k) * qpoint_stride + i_inner] = out[k];
                }
            }
          }
          for (int q = 0; q < nfq; ++q) coarse[v * nfq + q] += var_face[q];
        }
      }
    }
  }
};

// ---- artificial-viscosity generator -------------------------------------------------------------------------
// pde::Advection (pde.hpp:265-349) and pde::Smooth_art_visc (pde.hpp:355-431) through the Spatial kernels, their
// boundary hooks (Boundary_condition.cpp:12-60,346-367) and Solver::update_art_visc_smoothness / diffuse_art_visc
// (Solver.cpp:428-582). Element data: the Advection state is the momentum slots of `state` (holding velocity while
// the generator runs) followed by the rs `advection` arrays; Smooth_art_visc works on the 4 `forcing` arrays.
template <int nd, int rs>
struct Av
{
  static constexpr int nq = ipow(rs, nd), nfq = nq / rs, nv = nd + 2, nvert = 1 << nd;
  static constexpr int n_ext = nd + rs;  // Advection: n_extrap (velocity + rs advected scalars); n_update = rs
  static constexpr int n_sav = 3;        // Smooth_art_visc: n_extrap = n_update = 3
  using R = Rows<nd, rs>;
  using O = Ops<nd, rs>;

  static void element_faces(hx_host_mesh* m, const Mesh_view& mv, int cls, int e, double* (&faces)[6])
  {
    for (int f = 0; f < 2 * nd; ++f) faces[f] = mv.slot(m->face_slot[cls][e * 2 * nd + f]);
  }

  // Write_face (Spatial.hpp:41-57) of nvar gathered [nvar][nq] arrays
  template <int nvar>
  static void write_face_vars(const Basis& basis, const double (&ext)[nvar][nq], double* const* faces)
  {
    for (int i_dim = 0; i_dim < nd; ++i_dim)
      for (int i_fq = 0; i_fq < nfq; ++i_fq) {
        double row[rs][nvar];
        R::template read_row<nvar>(ext[0], i_dim, i_fq, row);
        double b[2][nvar];
        for (int s = 0; s < 2; ++s)
          for (int v = 0; v < nvar; ++v) {
            double acc = 0;
            for (int k = 0; k < rs; ++k) acc += basis.bound[s * rs + k] * row[k][v];
            b[s][v] = acc;
          }
        R::template write_bound<nvar>(b, faces, i_dim, i_fq);
      }
  }

  // Advection::fetch_extrap (pde.hpp:284-290)
  static void adv_gather(const hx_host_mesh* m, int cls, int e, double (&ext)[n_ext][nq])
  {
    const double* st = m->state[cls] + size_t(e) * nv * nq;
    const double* adv = m->advection[cls] + size_t(e) * rs * nq;
    for (int j = 0; j < nd; ++j)
      for (int q = 0; q < nq; ++q) ext[j][q] = st[j * nq + q];
    for (int i = 0; i < rs; ++i)
      for (int q = 0; q < nq; ++q) ext[nd + i][q] = adv[i * nq + q];
  }

  // Smooth_art_visc::fetch_extrap (pde.hpp:373-378): forcing[1..3]
  static void sav_gather(const hx_host_mesh* m, int cls, int e, double (&ext)[n_sav][nq])
  {
    const double* f = m->forcing[cls] + size_t(e) * 4 * nq;
    for (int i = 0; i < n_sav; ++i)
      for (int q = 0; q < nq; ++q) ext[i][q] = f[(1 + i) * nq + q];
  }

  static void adv_write_face(hx_host_mesh* m, int cls)
  {
    Basis basis(rs);
    Mesh_view mv{m};
    #pragma omp parallel for
    for (int e = 0; e < m->n_elem[cls]; ++e) {
      double ext[n_ext][nq];
      adv_gather(m, cls, e, ext);
      double* faces[6];
      element_faces(m, mv, cls, e, faces);
      write_face_vars<n_ext>(basis, ext, faces);
    }
  }

  static void sav_write_face(hx_host_mesh* m, int cls)
  {
    Basis basis(rs);
    Mesh_view mv{m};
    #pragma omp parallel for
    for (int e = 0; e < m->n_elem[cls]; ++e) {
      double ext[n_sav][nq];
      sav_gather(m, cls, e, ext);
      double* faces[6];
      element_faces(m, mv, cls, e, faces);
      write_face_vars<n_sav>(basis, ext, faces);
    }
  }

  // Advection nodes 2 x - 1 (pde.hpp:279-281)
  static void adv_nodes(const Basis& basis, double (&x)[rs])
  {
    for (int i = 0; i < rs; ++i) x[i] = 2 * basis.node[i] - 1.;
  }

  // Neighbor<Advection> (Spatial.hpp:613-704 with Advection::Computation, pde.hpp:319-347)
  static void adv_neighbor(hx_host_mesh* m, int cls)
  {
    const bool is_deformed = cls == HX_DEFORMED;
    Basis basis(rs);
    double x[rs];
    adv_nodes(basis, x);
    Mesh_view mv{m};
    #pragma omp parallel for
    for (int i_con = 0; i_con < m->n_con[cls]; ++i_con) {
      Dir dir = unpack_dir(m->con_dir[cls][i_con]);
      double* side[2] = {mv.slot(m->con_slot[cls][2 * i_con]), mv.slot(m->con_slot[cls][2 * i_con + 1])};
      double face[2][n_ext * nfq];
      double face_nrml[nd * nfq];
      int sign[2] = {1, 1};
      for (int s = 0; s < 2; ++s)
        for (int i = 0; i < n_ext * nfq; ++i) face[s][i] = side[s][i];
      Face_permutation<nd, rs> perm{dir, face[1], n_ext};
      if (is_deformed) {
        perm.match_faces();
        for (int s = 0; s < 2; ++s) sign[s] = 1 - 2 * dir.flip_normal(s);
        const double* nn = mv.nslot(m->con_nslot[2 * i_con]);
        for (int i = 0; i < nd * nfq; ++i) face_nrml[i] = nn[i];
      }
      for (int q = 0; q < nfq; ++q) {
        double n[nd];
        for (int j = 0; j < nd; ++j) n[j] = is_deformed ? sign[0] * face_nrml[j * nfq + q] : double(j == dir.i_dim[0]);
        double nvel[2], cs[2], a[2][rs];
        for (int s = 0; s < 2; ++s) {
          double vn = 0, vsq = 0;
          for (int j = 0; j < nd; ++j) {
            const double v = face[s][j * nfq + q];
            vn += v * n[j];
            vsq += v * v;
          }
          nvel[s] = vn;
          cs[s] = std::max(1., std::sqrt(vsq));
          for (int i = 0; i < rs; ++i) a[s][i] = face[s][(nd + i) * nfq + q];
        }
        double nsq = 0;
        for (int j = 0; j < nd; ++j) nsq += n[j] * n[j];
        const double coef = std::max(cs[0], cs[1]) * std::sqrt(nsq);
        for (int i = 0; i < rs; ++i) {
          const double f = .5 * (x[i] * nvel[0] * a[0][i] + x[i] * nvel[1] * a[1][i] + coef * (a[0][i] - a[1][i]));
          for (int s = 0; s < 2; ++s) face[s][i * nfq + q] = sign[s] * f;
        }
      }
      if (is_deformed) perm.restore();
      for (int s = 0; s < 2; ++s)
        for (int i = 0; i < rs * nfq; ++i) side[s][i] = face[s][i];
    }
  }

  // Neighbor<Smooth_art_visc>: diffusion only, the LDG average of the 3 extrapolated variables
  // Neighbor of a diffusion-only PDE (Smooth_art_visc: nvar 3; Fix_therm_admis: nvar nv): the LDG average only
  template <int nvar>
  static void diff_neighbor(hx_host_mesh* m, int cls)
  {
    const bool is_deformed = cls == HX_DEFORMED;
    Mesh_view mv{m};
    #pragma omp parallel for
    for (int i_con = 0; i_con < m->n_con[cls]; ++i_con) {
      Dir dir = unpack_dir(m->con_dir[cls][i_con]);
      double* side[2] = {mv.slot(m->con_slot[cls][2 * i_con]), mv.slot(m->con_slot[cls][2 * i_con + 1])};
      double face[4][nvar * nfq];
      for (int s = 0; s < 2; ++s)
        for (int i = 0; i < nvar * nfq; ++i) face[s][i] = side[s][i];
      Face_permutation<nd, rs> perm{dir, face[1], nvar};
      if (is_deformed) perm.match_faces();
      for (int q = 0; q < nfq; ++q)
        for (int v = 0; v < nvar; ++v) {
          const double avg = .5 * (face[0][v * nfq + q] + face[1][v * nfq + q]);
          for (int s = 0; s < 2; ++s) face[2 + s][v * nfq + q] = avg;
        }
      if (is_deformed) {
        perm.restore();
        Face_permutation<nd, rs>{dir, face[3], nvar}.restore();
      }
      for (int s = 0; s < 2; ++s)
        for (int i = 0; i < nvar * nfq; ++i) side[s][nv * nfq + i] = face[2 + s][i];
    }
  }
  static void sav_neighbor(hx_host_mesh* m, int cls) { diff_neighbor<n_sav>(m, cls); }

  // Neighbor_reconcile<Smooth_art_visc> (Spatial.hpp:716-759) on the 3 LDG variables
  static void sav_neighbor_reconcile(hx_host_mesh* m, int cls)
  {
    const bool is_deformed = cls == HX_DEFORMED;
    Mesh_view mv{m};
    #pragma omp parallel for
    for (int i_con = 0; i_con < m->n_con[cls]; ++i_con) {
      Dir dir = unpack_dir(m->con_dir[cls][i_con]);
      double* side[2] = {mv.slot(m->con_slot[cls][2 * i_con]) + nv * nfq,
                         mv.slot(m->con_slot[cls][2 * i_con + 1]) + nv * nfq};
      double face[2][n_sav * nfq];
      int sign[2] = {1, 1};
      for (int s = 0; s < 2; ++s)
        for (int i = 0; i < n_sav * nfq; ++i) face[s][i] = side[s][i];
      Face_permutation<nd, rs> perm{dir, face[1], n_sav};
      if (is_deformed) {
        perm.match_faces();
        for (int s = 0; s < 2; ++s) sign[s] = 1 - 2 * dir.flip_normal(s);
      }
      for (int q = 0; q < nfq; ++q)
        for (int v = 0; v < n_sav; ++v) {
          double avg = 0;
          for (int s = 0; s < 2; ++s) avg += .5 * sign[s] * face[s][v * nfq + q];
          for (int s = 0; s < 2; ++s) {
            double& f = face[s][v * nfq + q];
            f = sign[s] * avg - f;
          }
        }
      if (is_deformed) perm.restore();
      for (int s = 0; s < 2; ++s)
        for (int i = 0; i < n_sav * nfq; ++i) side[s][i] = face[s][i];
    }
  }

  // Local<Advection> (Spatial.hpp:326-509): convective rate, the constant source 2/advect_length in stage 0, and
  // Advection::write_update (pde.hpp:292-300), which divides by the implicit pseudo-time factor 1 + tss 2/L
  static void adv_local(hx_host_mesh* m, int cls, double dt, int stage, bool compute_residual, double advect_length)
  {
    const bool is_deformed = cls == HX_DEFORMED;
    Basis basis(rs);
    Derivative<rs> derivative(basis);
    double x[rs];
    adv_nodes(basis, x);
    const double update_coef = stage ? dt * basis.step_ratio() : dt;
    Mesh_view mv{m};
    #pragma omp parallel for
    for (int e = 0; e < m->n_elem[cls]; ++e) {
      double* state = m->state[cls] + size_t(e) * nv * nq;
      double* adv = m->advection[cls] + size_t(e) * rs * nq;
      double* faces[6];
      element_faces(m, mv, cls, e, faces);
      double* tss = m->tss[cls] + size_t(e) * nq;
      const double d_pos = m->nom_size[cls][e];
      const double* nrml = is_deformed ? m->metrics + size_t(e) * (nd * nd + 1) * nq : nullptr;
      const double* elem_det = is_deformed ? nrml + nd * nd * nq : nullptr;
      double time_rate[2][rs][nq] = {};
      double flux[nd][rs][nq];
      for (int q = 0; q < nq; ++q) {
        for (int i_dim = 0; i_dim < nd; ++i_dim) {
          double vn = 0;
          for (int j = 0; j < nd; ++j)
            vn += state[j * nq + q] * (is_deformed ? nrml[(i_dim * nd + j) * nq + q] : double(j == i_dim));
          for (int i = 0; i < rs; ++i) flux[i_dim][i][q] = x[i] * vn * adv[i * nq + q];
        }
        if (!stage) {
          double mult = d_pos;
          if (is_deformed) mult *= elem_det[q];
          for (int i = 0; i < rs; ++i) time_rate[1][i][q] = mult * (2 / advect_length);
        }
      }
      for (int i_dim = 0; i_dim < nd; ++i_dim)
        for (int i_fq = 0; i_fq < nfq; ++i_fq) {
          double row_f[rs][rs], face_f[2][rs], d[rs][rs];
          R::template read_row<rs>(flux[i_dim][0], i_dim, i_fq, row_f);
          R::template read_bound<rs>(faces, i_dim, i_fq, face_f);
          derivative.template apply<rs>(row_f, face_f, d);
          for (int k = 0; k < rs; ++k)
            for (int v = 0; v < rs; ++v) d[k][v] = -d[k][v];
          R::template write_row<rs>(d, time_rate[0][0], i_dim, i_fq, 1.);
        }
      double* ref_state = m->rcache[cls] + size_t(e) * std::max(nv, rs) * nq;
      for (int q = 0; q < nq; ++q) {
        double update[rs] = {};
        double mult = update_coef * tss[q] / d_pos;
        if (is_deformed) mult /= elem_det[q];
        for (int i = 0; i < rs; ++i) {
          double u = time_rate[0][i][q];
          if (stage) {
            u -= ref_state[i * nq + q];
          } else {
            ref_state[i * nq + q] = u;
            u += time_rate[1][i][q];
          }
          u *= mult;
          if (compute_residual) ref_state[i * nq + q] = u;
          else update[i] = u;
        }
        const double pseudo = 1 + tss[q] * 2 / advect_length;
        for (int i = 0; i < rs; ++i) {
          double& a = adv[i * nq + q];
          if (!stage) a = (a + update[i]) / pseudo;
          else a += update[i] / pseudo;
        }
      }
      double ext[n_ext][nq];
      adv_gather(m, cls, e, ext);
      write_face_vars<n_ext>(basis, ext, faces);
    }
  }

  // Local<Smooth_art_visc> (Spatial.hpp:326-509, stage 0): LDG gradient of forcing[1..3], flux -gradient*normal
  // (pde.hpp:405-410), the chained source of pde.hpp:419-427, and an additive update (write_update, pde.hpp:380-388)
  // Local (Spatial.hpp:326-509, stage 0) of a diffusion-only PDE on nvar contiguous variables: Smooth_art_visc
  // (forcing[1..3], with its source, pde.hpp:355-431) or Fix_therm_admis (the conserved state, pde.hpp:437-493)
  template <int n_sav, bool sav>
  static void diff_local(hx_host_mesh* m, int cls, double dt, double diff_time)
  {
    const bool is_deformed = cls == HX_DEFORMED;
    Basis basis(rs);
    Derivative<rs> derivative(basis);
    Mesh_view mv{m};
    static double cartesian_normal[nd][nd][nfq];
    for (int i = 0; i < nd; ++i)
      for (int j = 0; j < nd; ++j)
        for (int q = 0; q < nfq; ++q) cartesian_normal[i][j][q] = i == j;
    #pragma omp parallel for
    for (int e = 0; e < m->n_elem[cls]; ++e) {
      double* forcing = sav ? m->forcing[cls] + size_t(e) * 4 * nq : nullptr;
      double* vars = sav ? forcing + nq : m->state[cls] + size_t(e) * nv * nq;
      double* faces[6];
      double* visc_faces[6];
      element_faces(m, mv, cls, e, faces);
      for (int f = 0; f < 2 * nd; ++f) visc_faces[f] = faces[f] + nv * nfq;
      double* tss = m->tss[cls] + size_t(e) * nq;
      const double d_pos = m->nom_size[cls][e];
      const double* nrml = is_deformed ? m->metrics + size_t(e) * (nd * nd + 1) * nq : nullptr;
      const double* elem_det = is_deformed ? nrml + nd * nd * nq : nullptr;
      double* face_nrml[6] = {};
      if (is_deformed)
        for (int f = 0; f < 2 * nd; ++f) {
          int ns = m->face_nslot[e * 2 * nd + f];
          face_nrml[f] = ns >= 0 ? mv.nslot(ns) : cartesian_normal[f / 2][0];
        }
      double time_rate[2][n_sav][nq] = {};
      double visc_storage[nd][n_sav][nq] = {};
      for (int q = 0; q < nq; ++q)
        for (int v = 0; v < n_sav; ++v) time_rate[0][v][q] = vars[v * nq + q];
      for (int i_dim = 0; i_dim < nd; ++i_dim)
        for (int i_fq = 0; i_fq < nfq; ++i_fq) {
          double row_r[rs][n_sav], face_state[2][n_sav];
          R::template read_row<n_sav>(time_rate[0][0], i_dim, i_fq, row_r);
          R::template read_bound<n_sav>(visc_faces, i_dim, i_fq, face_state);
          if (is_deformed) {
            double row_n[rs][nd], face_n[2][nd];
            R::template read_row<nd>(nrml + i_dim * nd * nq, i_dim, i_fq, row_n);
            R::template read_bound<nd>(face_nrml, i_dim, i_fq, face_n);
            for (int j_dim = 0; j_dim < nd; ++j_dim)
              for (int v = 0; v < n_sav; ++v) {
                double row[rs][1], bound[2][1], d[rs][1];
                for (int k = 0; k < rs; ++k) row[k][0] = row_n[k][j_dim] * row_r[k][v];
                for (int s = 0; s < 2; ++s) bound[s][0] = face_n[s][j_dim] * face_state[s][v];
                derivative.template apply<1>(row, bound, d);
                for (int k = 0; k < rs; ++k) d[k][0] = d[k][0] / d_pos;
                R::template write_row<1>(d, visc_storage[j_dim][v], i_dim, i_fq, 1.);
              }
          } else {
            double d[rs][n_sav];
            derivative.template apply<n_sav>(row_r, face_state, d);
            for (int k = 0; k < rs; ++k)
              for (int v = 0; v < n_sav; ++v) d[k][v] = d[k][v] / d_pos;
            R::template write_row<n_sav>(d, visc_storage[i_dim][0], i_dim, i_fq, 0);
          }
        }
      std::memset(time_rate, 0, sizeof(time_rate));
      for (int q = 0; q < nq; ++q) {
        double grad[n_sav][nd];
        for (int i_dim = 0; i_dim < nd; ++i_dim)
          for (int v = 0; v < n_sav; ++v) grad[v][i_dim] = visc_storage[i_dim][v][q];
        if (is_deformed)
          for (int v = 0; v < n_sav; ++v)
            for (int i_dim = 0; i_dim < nd; ++i_dim) grad[v][i_dim] /= elem_det[q];
        for (int i_dim = 0; i_dim < nd; ++i_dim)
          for (int v = 0; v < n_sav; ++v) {
            double acc = 0;
            for (int j = 0; j < nd; ++j)
              acc += -grad[v][j] * (is_deformed ? nrml[(i_dim * nd + j) * nq + q] : double(j == i_dim));
            visc_storage[i_dim][v][q] = acc;
          }
        if (sav) {  // Smooth_art_visc::compute_source (pde.hpp:419-427)
          double mult = d_pos;
          if (is_deformed) mult *= elem_det[q];
          for (int v = 0; v < n_sav; ++v) {
            const double f = std::abs(forcing[v * nq + q]);
            time_rate[1][v][q] = mult * (((v == 1) ? std::sqrt(f) : f) / diff_time);
          }
        }
      }
      for (int i_dim = 0; i_dim < nd; ++i_dim)
        for (int i_fq = 0; i_fq < nfq; ++i_fq) {
          double row_f[rs][n_sav], face_f[2][n_sav], d[rs][n_sav];
          R::template read_row<n_sav>(visc_storage[i_dim][0], i_dim, i_fq, row_f);
          for (int s = 0; s < 2; ++s)
            for (int v = 0; v < n_sav; ++v) {
              double acc = 0;
              for (int k = 0; k < rs; ++k) acc += basis.bound[s * rs + k] * row_f[k][v];
              face_f[s][v] = acc;
            }
          for (int v = 0; v < n_sav; ++v)
            for (int s = 0; s < 2; ++s) visc_faces[i_dim * 2 + s][v * nfq + i_fq] = face_f[s][v];
          derivative.template apply<n_sav>(row_f, d);
          for (int k = 0; k < rs; ++k)
            for (int v = 0; v < n_sav; ++v) d[k][v] = -d[k][v];
          R::template write_row<n_sav>(d, time_rate[1][0], i_dim, i_fq, 1.);
        }
      for (int q = 0; q < nq; ++q) {
        double mult = dt * tss[q] / d_pos;
        if (is_deformed) mult /= elem_det[q];
        for (int v = 0; v < n_sav; ++v) {
          double u = time_rate[0][v][q];
          u += time_rate[1][v][q];
          u *= mult;
          vars[v * nq + q] += u;
        }
      }
    }
  }
  static void sav_local(hx_host_mesh* m, int cls, double dt, double diff_time) { diff_local<n_sav, true>(m, cls, dt, diff_time); }
  static void fta_local(hx_host_mesh* m, int cls, double dt) { diff_local<nv, false>(m, cls, dt, 1.); }

  // Reconcile_ldg_flux<Smooth_art_visc / Fix_therm_admis> (Spatial.hpp:543-594): lift the flux jump, update (for
  // Smooth_art_visc additionally divided by the pseudo-time factor), extrapolate to the faces. Smooth_art_visc:
  // division (critical write_update), extrapolate forcing[1..3] to the faces
  template <int n_sav, bool sav>
  static void diff_reconcile(hx_host_mesh* m, int cls, double dt, double diff_time, double cheby)
  {
    const bool is_deformed = cls == HX_DEFORMED;
    Basis basis(rs);
    Derivative<rs> derivative(basis);
    Mesh_view mv{m};
    #pragma omp parallel for
    for (int e = 0; e < m->n_elem[cls]; ++e) {
      double* vars = sav ? m->forcing[cls] + size_t(e) * 4 * nq + nq : m->state[cls] + size_t(e) * nv * nq;
      double* faces[6];
      double* visc_faces[6];
      element_faces(m, mv, cls, e, faces);
      for (int f = 0; f < 2 * nd; ++f) visc_faces[f] = faces[f] + nv * nfq;
      double* tss = m->tss[cls] + size_t(e) * nq;
      const double d_pos = m->nom_size[cls][e];
      const double* elem_det = is_deformed ? m->metrics + size_t(e) * (nd * nd + 1) * nq + nd * nd * nq : nullptr;
      double time_rate[n_sav][nq] = {};
      for (int i_dim = 0; i_dim < nd; ++i_dim)
        for (int i_fq = 0; i_fq < nfq; ++i_fq) {
          double bound_f[2][n_sav], d[rs][n_sav];
          R::template read_bound<n_sav>(visc_faces, i_dim, i_fq, bound_f);
          derivative.template boundary_term<n_sav>(bound_f, d);
          for (int k = 0; k < rs; ++k)
            for (int v = 0; v < n_sav; ++v) d[k][v] = -d[k][v];
          R::template write_row<n_sav>(d, time_rate[0], i_dim, i_fq, 1.);
        }
      for (int q = 0; q < nq; ++q) {
        double mult = dt * tss[q] / d_pos;
        if (is_deformed) mult /= elem_det[q];
        for (int v = 0; v < n_sav; ++v) {
          double& f = vars[v * nq + q];
          f += time_rate[v][q] * mult;
          if (sav) f /= 1 + tss[q] * cheby / diff_time;  // the critical write_update (pde.hpp:380-388)
        }
      }
      double ext[n_sav][nq];
      for (int v = 0; v < n_sav; ++v)
        for (int q = 0; q < nq; ++q) ext[v][q] = vars[v * nq + q];
      write_face_vars<n_sav>(basis, ext, faces);
    }
  }
  static void sav_reconcile(hx_host_mesh* m, int cls, double dt, double diff_time, double cheby)
  {
    diff_reconcile<n_sav, true>(m, cls, dt, diff_time, cheby);
  }
  static void fta_reconcile(hx_host_mesh* m, int cls, double dt) { diff_reconcile<nv, false>(m, cls, dt, 1., 0.); }

  // Max_dt (Spatial.hpp:784-828) for Advection (char_speed max(1, |v|), pde.hpp:341-345) or Smooth_art_visc
  // (diffusivity 1, pde.hpp:412-416)
  static double max_dt(hx_host_mesh* m, int cls, bool smooth, bool is_local, double safety_conv, double safety_diff)
  {
    Basis basis(rs);
    const double max_cfl_c = basis.max_cfl() * safety_conv;
    const double max_cfl_d = -2 / basis.min_eig_diffusion() * safety_diff;
    double dt = std::numeric_limits<double>::max();
    #pragma omp parallel for reduction(min : dt)
    for (int e = 0; e < m->n_elem[cls]; ++e) {
      const double* state = m->state[cls] + size_t(e) * nv * nq;
      double* tss = m->tss[cls] + size_t(e) * nq;
      for (int q = 0; q < nq; ++q) {
        double vs[nvert], coords[nd];
        for (int i = 0; i < nvert; ++i) vs[i] = m->vertex_tss[cls][e * nvert + i];
        for (int d = 0; d < nd; ++d) coords[d] = basis.node[(q / ipow(rs, nd - 1 - d)) % rs];
        const double spacing = O::interp(vs, coords);
        double scale = 0;
        if (smooth) {
          scale += 1. / max_cfl_d / spacing / spacing;
        } else {
          double vsq = 0;
          for (int j = 0; j < nd; ++j) vsq += state[j * nq + q] * state[j * nq + q];
          scale += std::max(1., std::sqrt(vsq)) / max_cfl_c / spacing;
        }
        if (is_local) tss[q] = 1. / scale;
        else {
          tss[q] = 1.;
          dt = std::min(dt, 1. / scale);
        }
      }
    }
    return is_local ? 1. : dt;
  }
};

// boundary hooks of the generator on the deformed-class boundary connections
void apply_advection_bcs(hx_host_mesh* m)
{
  Mesh_view mv{m};
  const int nd = m->n_dim, rs = m->row_size, nv = nd + 2, nfq = ipow(rs, nd - 1);
  #pragma omp parallel for
  for (int i = 0; i < m->n_bc_con; ++i) {
    int con = m->bc_con[i];
    const hx_bc& bc = m->bcs[m->bc_serial[i]];
    double* in_f = mv.slot(m->con_slot[HX_DEFORMED][2 * con]);
    double* gh_f = mv.slot(m->con_slot[HX_DEFORMED][2 * con + 1]);
    const double* nrml = mv.nslot(m->con_nslot[2 * con]);
    switch (bc.kind) {
      case HX_BC_COPY:  // Copy::apply_advection = copy_state (Boundary_condition.cpp:12-22,460-463)
        for (int l = 0; l < 2; ++l)
          for (int k = 0; k < nv * nfq; ++k) gh_f[l * nv * nfq + k] = in_f[l * nv * nfq + k];
        break;
      case HX_BC_NONPENETRATION:  // Boundary_condition.cpp:346-367
        for (int q = 0; q < nfq; ++q) {
          double dot = 0., nsq = 0.;
          for (int d = 0; d < nd; ++d) {
            dot += in_f[d * nfq + q] * nrml[d * nfq + q];
            nsq += nrml[d * nfq + q] * nrml[d * nfq + q];
          }
          for (int d = 0; d < nd; ++d) gh_f[d * nfq + q] = in_f[d * nfq + q] - 2 * dot * nrml[d * nfq + q] / nsq;
          for (int a = 0; a < rs; ++a) gh_f[(nd + a) * nfq + q] = in_f[(nd + a) * nfq + q];
        }
        break;
      case HX_BC_NO_SLIP:  // Boundary_condition.cpp:429-448
        for (int k = 0; k < nd * nfq; ++k) gh_f[k] = -in_f[k];
        for (int k = 0; k < rs * nfq; ++k) gh_f[nd * nfq + k] = in_f[nd * nfq + k];
        break;
      default:  // Flow_bc::apply_advection (Boundary_condition.cpp:24-41)
        for (int k = 0; k < nd * nfq; ++k) gh_f[k] = in_f[k];
        for (int k = 0; k < rs * nfq; ++k) gh_f[nd * nfq + k] = 2. - in_f[nd * nfq + k];
        break;
    }
  }
}

// Solver::apply_avc_diff_bcs / apply_avc_diff_flux_bcs (Solver.cpp:83-101) with Flow_bc::apply_diffusion and
// flux_diffusion (Boundary_condition.cpp:43-60), which no condition built here overrides
void apply_avc_diff_bcs(hx_host_mesh* m, bool flux)
{
  Mesh_view mv{m};
  const int nd = m->n_dim, nv = nd + 2, nfq = ipow(m->row_size, nd - 1);
  #pragma omp parallel for
  for (int i = 0; i < m->n_bc_con; ++i) {
    int con = m->bc_con[i];
    double* in_f = mv.slot(m->con_slot[HX_DEFORMED][2 * con]);
    double* gh_f = mv.slot(m->con_slot[HX_DEFORMED][2 * con + 1]);
    if (flux)
      for (int k = 0; k < nv * nfq; ++k) gh_f[nv * nfq + k] = -in_f[nv * nfq + k];
    else
      for (int k = 0; k < nv * nfq; ++k) gh_f[k] = in_f[k];
  }
}

// ---- dispatch ---------------------------------------------------------------------------------------------

template <template <int, int> class F, typename... A>
auto dispatch(int nd, int rs, A&&... args)
{
  switch (nd * 16 + rs) {
#define HX_CASE(ND, RS) \
  case ND * 16 + RS: return F<ND, RS>::run(std::forward<A>(args)...);
#define HX_CASES(ND) HX_CASE(ND, 2) HX_CASE(ND, 3) HX_CASE(ND, 4) HX_CASE(ND, 5) HX_CASE(ND, 6) HX_CASE(ND, 7) HX_CASE(ND, 8)
    HX_CASES(1) HX_CASES(2) HX_CASES(3)
#undef HX_CASES
#undef HX_CASE
  }
  throw std::runtime_error("demand for invalid kernel");
}

void check_mesh(const hx_host_mesh* m)
{
  if (!m || m->n_dim < 1 || m->n_dim > 3 || m->row_size < 2 || m->row_size > 8)
    throw std::runtime_error("demand for invalid kernel");
}

bool is_ns(const hx_pde* pde)
{
  if (!pde || pde->kind == HX_PDE_EULER) return false;
  if (pde->kind == HX_PDE_NAVIER_STOKES) return true;
  throw std::runtime_error("oracle: PDE kind not restated");
}

template <int nd, bool visc>
Navier_stokes<nd, visc> make_eq(const hx_pde* pde)
{
  Navier_stokes<nd, visc> eq;
  if (pde && pde->kind == HX_PDE_NAVIER_STOKES) {
    eq.dyn_visc = Transport(pde->visc);
    eq.therm_cond = Transport(pde->cond);
  }
  return eq;
}

template <int nd, int rs>
struct Run_write_face { static int run(hx_host_mesh* m, int cls) { Ops<nd, rs>::write_face(m, cls); return 0; } };

template <int nd, int rs>
struct Run_local
{
  static int run(hx_host_mesh* m, int cls, const hx_pde* pde, double dt, int stage, int cr, int uf)
  {
    if (is_ns(pde)) Ops<nd, rs>::template local<true>(m, cls, make_eq<nd, true>(pde), dt, stage, cr, uf);
    else Ops<nd, rs>::template local<false>(m, cls, make_eq<nd, false>(pde), dt, stage, cr, uf);
    return 0;
  }
};

template <int nd, int rs>
struct Run_neighbor
{
  static int run(hx_host_mesh* m, int cls, const hx_pde* pde)
  {
    if (is_ns(pde)) Ops<nd, rs>::template neighbor<true>(m, cls, make_eq<nd, true>(pde));
    else Ops<nd, rs>::template neighbor<false>(m, cls, make_eq<nd, false>(pde));
    return 0;
  }
};

template <int nd, int rs>
struct Run_neighbor_reconcile { static int run(hx_host_mesh* m, int cls) { Ops<nd, rs>::neighbor_reconcile(m, cls); return 0; } };

template <int nd, int rs>
struct Run_reconcile
{
  static int run(hx_host_mesh* m, int cls, double dt, int cr, int uf)
  {
    Ops<nd, rs>::template reconcile_ldg_flux<true>(m, cls, dt, cr, uf);
    return 0;
  }
};

template <int nd, int rs>
struct Run_max_dt
{
  static double run(hx_host_mesh* m, int cls, const hx_pde* pde, int is_local, double sc, double sd)
  {
    if (is_ns(pde)) return Ops<nd, rs>::template max_dt<true>(m, cls, make_eq<nd, true>(pde), is_local, sc, sd);
    return Ops<nd, rs>::template max_dt<false>(m, cls, make_eq<nd, false>(pde), is_local, sc, sd);
  }
};

template <int nd, int rs>
struct Run_prolong { static int run(hx_host_mesh* m, int nvar, int s, int o) { Ops<nd, rs>::prolong(m, nvar, s, o); return 0; } };
template <int nd, int rs>
struct Run_restrict { static int run(hx_host_mesh* m, int nvar, int s, int o) { Ops<nd, rs>::restrict_(m, nvar, s, o); return 0; } };

int pde_kind(const hx_pde* pde) { return pde ? pde->kind : HX_PDE_EULER; }

int n_extrap(const hx_host_mesh* m, const hx_pde* pde)
{
  int kind = pde_kind(pde);
  if (kind == HX_PDE_EULER || kind == HX_PDE_NAVIER_STOKES) return m->n_dim + 2;
  if (kind == HX_PDE_ADVECTION) return m->n_dim + m->row_size;
  if (kind == HX_PDE_SMOOTH_AV) return 3;
  if (kind == HX_PDE_FIX_ADMIS) return m->n_dim + 2;
  throw std::runtime_error("oracle: PDE kind not restated");
}

template <int nd, int rs>
struct Run_av_write_face
{
  static int run(hx_host_mesh* m, int cls, int kind)
  {
    if (kind == HX_PDE_ADVECTION) Av<nd, rs>::adv_write_face(m, cls);
    else if (kind == HX_PDE_SMOOTH_AV) Av<nd, rs>::sav_write_face(m, cls);
    else Ops<nd, rs>::write_face(m, cls);  // Write_face<Fix_therm_admis>: the nv state variables
    return 0;
  }
};
template <int nd, int rs>
struct Run_av_neighbor
{
  static int run(hx_host_mesh* m, int cls, int kind)
  {
    if (kind == HX_PDE_ADVECTION) Av<nd, rs>::adv_neighbor(m, cls);
    else if (kind == HX_PDE_SMOOTH_AV) Av<nd, rs>::sav_neighbor(m, cls);
    else Av<nd, rs>::template diff_neighbor<nd + 2>(m, cls);
    return 0;
  }
};
template <int nd, int rs>
struct Run_adv_local
{
  static int run(hx_host_mesh* m, int cls, double dt, int stage, int cr, double advect_length)
  {
    Av<nd, rs>::adv_local(m, cls, dt, stage, cr, advect_length);
    return 0;
  }
};
template <int nd, int rs>
struct Run_sav_local
{
  static int run(hx_host_mesh* m, int cls, double dt, double diff_time) { Av<nd, rs>::sav_local(m, cls, dt, diff_time); return 0; }
};
template <int nd, int rs>
struct Run_sav_neighbor_reconcile { static int run(hx_host_mesh* m, int cls) { Av<nd, rs>::sav_neighbor_reconcile(m, cls); return 0; } };
template <int nd, int rs>
struct Run_sav_reconcile
{
  static int run(hx_host_mesh* m, int cls, double dt, double diff_time, double cheby)
  {
    Av<nd, rs>::sav_reconcile(m, cls, dt, diff_time, cheby);
    return 0;
  }
};
template <int nd, int rs>
struct Run_face_permutation
{
  static int run(Dir dir, double* data, int n_var, bool match)
  {
    Face_permutation<nd, rs> fp{dir, data, n_var};
    if (match) fp.match_faces();
    else fp.restore();
    return 0;
  }
};
template <int nd, int rs>
struct Run_fta_local { static int run(hx_host_mesh* m, int cls, double dt) { Av<nd, rs>::fta_local(m, cls, dt); return 0; } };
template <int nd, int rs>
struct Run_fta_reconcile { static int run(hx_host_mesh* m, int cls, double dt) { Av<nd, rs>::fta_reconcile(m, cls, dt); return 0; } };
template <int nd, int rs>
struct Run_av_max_dt
{
  static double run(hx_host_mesh* m, int cls, bool smooth, int is_local, double sc, double sd)
  {
    return Av<nd, rs>::max_dt(m, cls, smooth, is_local, sc, sd);
  }
};

// the Solver-level point loops of update_art_visc_smoothness (Solver.cpp:465-477,525-540,549-574)
template <int nd, int rs>
struct Run_av_points
{
  static constexpr int nq = ipow(rs, nd), nv = nd + 2;
  // stage 0: momentum /= sqrt(2 rho E); stage 1: Legendre projection of the advected scalars -> forcing[0];
  // stage 2: new coefficient, residual and momentum restored
  static double run(hx_host_mesh* m, int what, const hx_av_options* o)
  {
    Basis basis(rs);
    double resid = 0;
    const int n_real = 3;  // Storage_params::n_forcing - 1
    const double mult = o->visc_mult * o->advect_length;
    const double us_max = o->advect_length * o->unscaled_max * std::sqrt(2 * o->freestream_energy / o->freestream_mass);
    double qw[nq];
    for (int q = 0; q < nq; ++q) {
      double w = 1;
      for (int d = 0; d < nd; ++d) w *= basis.weight[(q / ipow(rs, nd - 1 - d)) % rs];
      qw[q] = w;
    }
    for (int cls = 0; cls < 2; ++cls) {
      #pragma omp parallel for reduction(+ : resid)
      for (int e = 0; e < m->n_elem[cls]; ++e) {
        double* state = m->state[cls] + size_t(e) * nv * nq;
        if (what == 0) {
          for (int q = 0; q < nq; ++q) {
            double scale = std::sqrt(2 * state[nd * nq + q] * state[(nd + 1) * nq + q]);
            for (int d = 0; d < nd; ++d) state[d * nq + q] /= scale;
          }
        } else if (what == 1) {
          const double* adv = m->advection[cls] + size_t(e) * rs * nq;
          double* forcing = m->forcing[cls] + size_t(e) * 4 * nq;
          const double* orth = basis.ortho + (o->av_rs - 1) * rs;
          for (int q = 0; q < nq; ++q) {
            double proj = 0;
            for (int i = 0; i < rs; ++i) proj += adv[i * nq + q] * basis.weight[i] * orth[i];
            forcing[q] = proj * proj * 2 * state[(nd + 1) * nq + q] / state[nd * nq + q];
          }
        } else {
          double* av = m->av[cls] + size_t(e) * 2 * nq;
          const double* forcing = m->forcing[cls] + size_t(e) * 4 * nq;
          const double volume = mpow(m->nom_size[cls][e], nd);
          for (int q = 0; q < nq; ++q) {
            const double f = mult * forcing[n_real * nq + q];
            const double new_av = us_max * f / (us_max + f);
            resid += mpow(av[q] - new_av, 2) * qw[q] * volume;
            av[q] = new_av;
            double scale = std::sqrt(2 * state[nd * nq + q] * state[(nd + 1) * nq + q]);
            for (int d = 0; d < nd; ++d) state[d * nq + q] *= scale;
          }
        }
      }
    }
    return resid;
  }
};

// Characteristics of the Euler flux about `ref` in direction `n` (pde.hpp:182-260), for Riemann_invariants. The
// 3x3 system of the 1D decomposition is solved by Cramer's rule (the reference uses column-pivoting QR; the systems
// here are nonsingular whenever the pressure is positive).
struct Char_decomp
{
  int nd;
  double vals[3], vecs[3][3], dir[3], veloc[3];
  Char_decomp(int nd_, const double* ref, const double* n) : nd{nd_}
  {
    double norm = 0.;
    for (int d = 0; d < nd; ++d) norm += n[d] * n[d];
    norm = std::sqrt(norm);
    double vsq = 0., vn = 0.;
    for (int d = 0; d < nd; ++d) {
      dir[d] = n[d] / norm;
      veloc[d] = ref[d] / ref[nd];
    }
    for (int d = 0; d < nd; ++d) {
      vsq += veloc[d] * veloc[d];
      vn += dir[d] * veloc[d];
    }
    const double mass = ref[nd];
    const double pres = .4 * (ref[nd + 1] - .5 * mass * vsq);
    const double c = std::sqrt(1.4 * std::max(pres, 0.) / mass);
    vals[0] = vn - c;
    vals[1] = vn + c;
    vals[2] = vn;
    for (int k = 0; k < 2; ++k) {
      const double dv = (2 * k - 1) * c / mass;
      vecs[0][k] = vn + mass * dv;
      vecs[1][k] = 1.;
      vecs[2][k] = 1.4 * pres / mass / .4 + .5 * vsq + mass * vn * dv;
    }
    vecs[0][2] = vn;
    vecs[1][2] = 1.;
    vecs[2][2] = .5 * vsq;
  }
  static double det3(const double a[3][3])
  {
    return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
           a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  }
  // part[v][c]: the component of `u` in the eigenspace of vals[c]; the parts sum to u
  void split(const double* u, double part[5][3]) const
  {
    double un = 0., vn = 0.;
    for (int d = 0; d < nd; ++d) {
      un += dir[d] * u[d];
      vn += dir[d] * veloc[d];
    }
    double vt[3], corr[3], vc = 0.;
    for (int d = 0; d < nd; ++d) {
      vt[d] = veloc[d] - vn * dir[d];
      corr[d] = u[d] - un * dir[d] - u[nd] * vt[d];
      vc += veloc[d] * corr[d];
    }
    const double rhs[3] = {un, u[nd], u[nd + 1] - vc};
    const double det = det3(vecs);
    if (det == 0.) throw std::runtime_error("singular characteristic system");
    double coef[3];
    for (int c = 0; c < 3; ++c) {
      double a[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[i][j] = j == c ? rhs[i] : vecs[i][j];
      coef[c] = det3(a) / det;
    }
    for (int c = 0; c < 3; ++c) {
      for (int d = 0; d < nd; ++d) part[d][c] = dir[d] * vecs[0][c] * coef[c] + vt[d] * coef[c];
      part[nd][c] = vecs[1][c] * coef[c];
      part[nd + 1][c] = vecs[2][c] * coef[c];
    }
    for (int d = 0; d < nd; ++d) part[d][2] += corr[d];
    part[nd + 1][2] += vc;
  }
  // apply_char (Boundary_condition.cpp:83-95)
  void choose(int sign, const double* inside, const double* outside, double* out) const
  {
    double pi[5][3], po[5][3];
    split(inside, pi);
    split(outside, po);
    for (int v = 0; v < nd + 2; ++v) {
      out[v] = 0.;
      for (int c = 0; c < 3; ++c) out[v] += sign * vals[c] > 0 ? po[v][c] : pi[v][c];
    }
  }
};

// No_slip's thermal condition (Boundary_condition.hpp:140-186, Boundary_condition.cpp:373-379)
constexpr double sigma_sb()
{
  double pi5 = 1., k4 = 1., c2 = 1., h3 = 1.;
  for (int i = 0; i < 5; ++i) pi5 *= M_PI;
  for (int i = 0; i < 4; ++i) k4 *= 1.380649e-23;
  for (int i = 0; i < 2; ++i) c2 *= 299792458.;
  for (int i = 0; i < 3; ++i) h3 *= 6.62607015e-34;
  return 2 * pi5 * k4 / (15 * c2 * h3);  // constants.hpp:52
}
double wall_energy(const hx_bc& bc, const double* u, int nv)
{
  return bc.thermal == HX_THERMAL_ENERGY ? bc.thermal_value * u[nv - 2] : u[nv - 1];
}
double wall_heat_flux(const hx_bc& bc, const double* u, int nv, double q)
{
  switch (bc.thermal) {
    case HX_THERMAL_ENERGY: return q;
    case HX_THERMAL_HEAT_FLUX: return bc.thermal_value;
    default: {
      const double temp = u[nv - 1] * .4 / u[nv - 2] / 287.05287;
      return bc.emissivity * sigma_sb() * (temp * temp * temp * temp) + bc.heat_transfer_coef * (temp - bc.temperature);
    }
  }
}

bool needs_cache(int kind) { return kind == HX_BC_RIEMANN_INVARIANTS || kind == HX_BC_NO_SLIP; }

// Boundary conditions on the deformed-class boundary connections (Boundary_condition.cpp)
void apply_state_bcs(hx_host_mesh* m)
{
  for (int i = 0; i < m->n_bc; ++i)
    if (needs_cache(m->bcs[i].kind) && m->n_bc_con && !m->bc_cache)
      throw std::runtime_error("Riemann_invariants and No_slip need hx_host_mesh::bc_cache");
  Mesh_view mv{m};
  const int nd = m->n_dim, nv = nd + 2, nfq = ipow(m->row_size, nd - 1);
  #pragma omp parallel for
  for (int i = 0; i < m->n_bc_con; ++i) {
    int con = m->bc_con[i];
    const hx_bc& bc = m->bcs[m->bc_serial[i]];
    double* in_f = mv.slot(m->con_slot[HX_DEFORMED][2 * con]);
    double* gh_f = mv.slot(m->con_slot[HX_DEFORMED][2 * con + 1]);
    const double* nrml = mv.nslot(m->con_nslot[2 * con]);
    switch (bc.kind) {
      case HX_BC_FREESTREAM:  // Boundary_condition.cpp:66-76
        for (int v = 0; v < nv; ++v)
          for (int q = 0; q < nfq; ++q) gh_f[v * nfq + q] = bc.state[v];
        break;
      case HX_BC_COPY:     // copy_state (Boundary_condition.cpp:12-22): both halves
      case HX_BC_OUTFLOW:
        for (int l = 0; l < 2; ++l)
          for (int k = 0; k < nv * nfq; ++k) gh_f[l * nv * nfq + k] = in_f[l * nv * nfq + k];
        break;
      case HX_BC_NONPENETRATION:  // reflect_momentum (Boundary_condition.cpp:302-326)
        for (int k = 0; k < nv * nfq; ++k) gh_f[k] = in_f[k];
        for (int q = 0; q < nfq; ++q) {
          double dot = 0., norm_sq = 0.;
          for (int d = 0; d < nd; ++d) {
            double n = nrml[d * nfq + q];
            dot += gh_f[d * nfq + q] * n;
            norm_sq += n * n;
          }
          for (int d = 0; d < nd; ++d) gh_f[d * nfq + q] -= 2 * dot * nrml[d * nfq + q] / norm_sq;
        }
        break;
      case HX_BC_RIEMANN_INVARIANTS:
      case HX_BC_PRESSURE_OUTFLOW:
      case HX_BC_NO_SLIP: {
        const int face_sign = (m->con_dir[HX_DEFORMED][con] >> 4) & 1;  // inside_face_sign
        double* sc = m->bc_cache + size_t(i) * nv * nfq;
        for (int q = 0; q < nfq; ++q) {
          double in[5], n[3], g[5];
          for (int v = 0; v < nv; ++v) in[v] = in_f[v * nfq + q];
          for (int d = 0; d < nd; ++d) n[d] = nrml[d * nfq + q];
          if (bc.kind == HX_BC_RIEMANN_INVARIANTS) {  // Boundary_condition.cpp:97-141
            Char_decomp(nd, in, n).choose(1 - 2 * face_sign, in, bc.state, g);
            g[nd] = std::max(g[nd], in[nd] / 2);
            double ke = 0., ike = 0.;
            for (int d = 0; d < nd; ++d) {
              ke += g[d] * g[d];
              ike += in[d] * in[d];
            }
            ke = .5 * ke / g[nd];
            ike = .5 * ike / in[nd];
            g[nd + 1] = std::max(ke + std::max(g[nd + 1] - ke, (in[nd + 1] - ike) / 2), 0.);
            for (int v = 0; v < nv; ++v) sc[v * nfq + q] = in[v];
          } else if (bc.kind == HX_BC_PRESSURE_OUTFLOW) {  // Boundary_condition.cpp:184-214
            double mn = 0., nn = 0., mm = 0.;
            for (int d = 0; d < nd; ++d) {
              mn += in[d] * n[d];
              nn += n[d] * n[d];
              mm += in[d] * in[d];
            }
            const double ke = .5 * mm / in[nd];
            const double p = std::max(.4 * (in[nd + 1] - ke), 0.);
            for (int v = 0; v < nv; ++v) g[v] = in[v];
            if (mn / in[nd] / std::sqrt(nn) * (2 * face_sign - 1) < std::sqrt(1.4 * p / in[nd])) g[nd + 1] = bc.pressure / .4 + ke;
          } else {  // Boundary_condition.cpp:381-400
            for (int d = 0; d < nd; ++d) g[d] = -in[d];
            g[nd] = in[nd];
            const double e = wall_energy(bc, in, nv);
            g[nd + 1] = e * e / in[nd + 1];
            for (int v = 0; v < nv; ++v) sc[v * nfq + q] = (g[v] + in[v]) / 2;
          }
          for (int v = 0; v < nv; ++v) gh_f[v * nfq + q] = g[v];
        }
        break;
      }
      default:
        break;
    }
  }
}

void apply_flux_bcs(hx_host_mesh* m)
{
  for (int i = 0; i < m->n_bc; ++i)
    if (needs_cache(m->bcs[i].kind) && m->n_bc_con && !m->bc_cache)
      throw std::runtime_error("Riemann_invariants and No_slip need hx_host_mesh::bc_cache");
  Mesh_view mv{m};
  const int nd = m->n_dim, nv = nd + 2, nfq = ipow(m->row_size, nd - 1);
  #pragma omp parallel for
  for (int i = 0; i < m->n_bc_con; ++i) {
    int con = m->bc_con[i];
    const hx_bc& bc = m->bcs[m->bc_serial[i]];
    double* in_f = mv.slot(m->con_slot[HX_DEFORMED][2 * con]);
    double* gh_f = mv.slot(m->con_slot[HX_DEFORMED][2 * con + 1]);
    const double* nrml = mv.nslot(m->con_nslot[2 * con]);
    switch (bc.kind) {
      case HX_BC_FREESTREAM:  // Freestream::apply_flux = copy_state (Boundary_condition.cpp:299)
      case HX_BC_COPY:
        for (int l = 0; l < 2; ++l)
          for (int k = 0; k < nv * nfq; ++k) gh_f[l * nv * nfq + k] = in_f[l * nv * nfq + k];
        break;
      case HX_BC_OUTFLOW:  // Boundary_condition.cpp:470-477
        for (int k = 0; k < nv * nfq; ++k) gh_f[nv * nfq + k] = -in_f[nv * nfq + k];
        break;
      case HX_BC_NONPENETRATION: {  // Boundary_condition.cpp:333-345
        double* g = gh_f + nv * nfq;
        const double* f = in_f + nv * nfq;
        for (int k = 0; k < nv * nfq; ++k) g[k] = -f[k];
        for (int q = 0; q < nfq; ++q) {
          double dot = 0., norm_sq = 0.;
          for (int d = 0; d < nd; ++d) {
            double n = nrml[d * nfq + q];
            dot += g[d * nfq + q] * n;
            norm_sq += n * n;
          }
          for (int d = 0; d < nd; ++d) g[d * nfq + q] -= 2 * dot * nrml[d * nfq + q] / norm_sq;
        }
        break;
      }
      case HX_BC_PRESSURE_OUTFLOW:  // Boundary_condition.cpp:216-223
        for (int k = 0; k < nv * nfq; ++k) gh_f[nv * nfq + k] = -in_f[nv * nfq + k];
        break;
      case HX_BC_RIEMANN_INVARIANTS:
      case HX_BC_NO_SLIP: {
        const int face_sign = (m->con_dir[HX_DEFORMED][con] >> 4) & 1;
        const double* sc = m->bc_cache + size_t(i) * nv * nfq;
        const double* f = in_f + nv * nfq;
        double* g = gh_f + nv * nfq;
        for (int q = 0; q < nfq; ++q) {
          double fl[5], cs[5], n[3];
          for (int v = 0; v < nv; ++v) {
            fl[v] = f[v * nfq + q];
            cs[v] = sc[v * nfq + q];
          }
          for (int d = 0; d < nd; ++d) n[d] = nrml[d * nfq + q];
          if (bc.kind == HX_BC_RIEMANN_INVARIANTS) {  // Boundary_condition.cpp:143-182
            const double zero[5] = {0., 0., 0., 0., 0.};
            double out[5];
            Char_decomp(nd, cs, n).choose(1 - 2 * face_sign, zero, fl, out);
            for (int v = 0; v < nv; ++v) g[v * nfq + q] = out[v];
          } else {  // Boundary_condition.cpp:402-427
            for (int d = 0; d < nd; ++d) g[d * nfq + q] = fl[d];
            g[nd * nfq + q] = -fl[nd];
            double nn = 0.;
            for (int d = 0; d < nd; ++d) nn += n[d] * n[d];
            const double normal = std::sqrt(nn);
            const int fs = 2 * face_sign - 1;
            const double qe = fl[nd + 1];
            g[(nd + 1) * nfq + q] = bc.coercion * (normal * fs * wall_heat_flux(bc, cs, nv, qe * fs / normal) - qe) + qe;
          }
        }
        break;
      }
      default:
        break;
    }
  }
}

template <typename Fn>
int guarded(Fn&& fn)
{
  try {
    fn();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

hx_pde euler_pde()
{
  hx_pde p{};
  p.kind = HX_PDE_EULER;
  return p;
}

hx_pde ns_pde(hx_transport visc, hx_transport cond)
{
  hx_pde p{};
  p.kind = HX_PDE_NAVIER_STOKES;
  p.visc = visc;
  p.cond = cond;
  return p;
}

// ---- thermodynamic admissibility (Solver.cpp:921-1095) -----------------------------------------------------------

// thermo::admissible (thermo.cpp:6-18): density and energy positive at every point; a value that is not finite throws
bool admissible(const double* data, int nd, int n_qpoint)
{
  bool admiss = true;
  for (int q = 0; q < n_qpoint; ++q) {
    admiss = admiss && (data[nd * n_qpoint + q] > 0.) && (data[(nd + 1) * n_qpoint + q] > 0.);
    for (int v = 0; v < nd + 2; ++v)
      if (!std::isfinite(data[v * n_qpoint + q])) throw std::runtime_error("state is not finite");
  }
  return admiss;
}

// Solver::is_admissible (Solver.cpp:921-959); record[E] over the elements of both classes, Cartesian first. Per element
// the faces are checked in order until one is inadmissible, as the reference's && does. Every fine face of the hanging
// faces is checked (the reference's short-circuit over them depends on how OpenMP splits that loop).
bool is_admissible(hx_host_mesh* m, std::vector<int>& record)
{
  const int nd = m->n_dim, nq = ipow(m->row_size, nd), nfq = nq / m->row_size;
  Mesh_view mv{m};
  record.assign(size_t(m->n_elem[0]) + m->n_elem[1], 0);
  bool admiss = true;
  for (int c = 0; c < 2; ++c)
    for (int e = 0; e < m->n_elem[c]; ++e) {
      bool elem_admis = admissible(m->state[c] + size_t(e) * (nd + 2) * nq, nd, nq);
      for (int f = 0; f < 2 * nd; ++f)
        elem_admis = elem_admis && admissible(mv.slot(m->face_slot[c][e * 2 * nd + f]), nd, nfq);
      if (!elem_admis) record[(c ? m->n_elem[0] : 0) + e] = 1;
      admiss = admiss && elem_admis;
    }
  bool refined_admiss = true;
  for (int i = 0; i < m->n_ref; ++i) {
    int n_fine = 1 << (nd - 1);
    for (int d = 0; d < nd - 1; ++d)
      if (m->ref_stretch[2 * i + d]) n_fine /= 2;
    for (int f = 0; f < n_fine; ++f) {
      const bool a = admissible(mv.slot(m->ref_slot[5 * i + 1 + f]), nd, nfq);
      refined_admiss = refined_admiss && a;
    }
  }
  return admiss && refined_admiss;
}

// Vertex connectivity, which hx_host_mesh does not carry: the shared vertex of every element vertex and the
// hanging-vertex matchers of the hanging faces (set by oracle_set_vertices)
struct Vertex_data
{
  bool set = false;
  int n_elem[2] = {0, 0};
  int n_vertex = 0;
  std::vector<int> vid;      // [n_elem[0] + n_elem[1]][nvert]
  std::vector<int> hanging;  // [n][12]: i_dim, is_positive, stretch[2], then (class, element) of up to 4 fine elements
};
Vertex_data g_vertices;

// math::hypercube_matvec (math.cpp:19-56): `mat` (rows x cols, row-major) applied along every dimension of a hypercubic
// array, the last dimension first
std::vector<double> hypercube_matvec(const std::vector<double>& mat, int rows, int cols, const std::vector<double>& vec)
{
  const int vec_row_size = int(vec.size()) / cols;
  if (vec_row_size == 1) {
    std::vector<double> out(rows);
    for (int r = 0; r < rows; ++r) {
      double acc = 0.;
      for (int c = 0; c < cols; ++c) acc += mat[r * cols + c] * vec[c];
      out[r] = acc;
    }
    return out;
  }
  int prod_size = 1;
  for (int vs = int(vec.size()); vs > 1;) {
    prod_size *= rows;
    vs /= cols;
  }
  const int row_size = prod_size / rows;
  std::vector<double> fact(size_t(cols) * row_size);
  for (int i_row = 0; i_row < cols; ++i_row) {
    std::vector<double> sub(vec.begin() + i_row * vec_row_size, vec.begin() + (i_row + 1) * vec_row_size);
    std::vector<double> r = hypercube_matvec(mat, rows, cols, sub);
    std::copy(r.begin(), r.end(), fact.begin() + i_row * row_size);
  }
  std::vector<double> prod(prod_size);
  for (int i_col = 0; i_col < row_size; ++i_col)
    for (int r = 0; r < rows; ++r) {
      double acc = 0.;
      for (int c = 0; c < cols; ++c) acc += mat[r * cols + c] * fact[c * row_size + i_col];
      prod[r * row_size + i_col] = acc;
    }
  return prod;
}

// math::stretched_ind (math.hpp:188-199)
int stretched_ind(int nd, int ind, const int* stretch)
{
  int stride = 1, stretched = 0;
  for (int i_dim = nd - 2; i_dim >= 0; --i_dim)
    if (!stretch[i_dim]) {
      stretched += ((ind / ipow(2, nd - 2 - i_dim)) % 2) * stride;
      stride *= 2;
    }
  return stretched;
}

// Solver::share_vertex_data with Vertex::vector_max or vector_min (Solver.cpp:40-54, Vertex.cpp:6-14): every
// element's copy of a vertex value becomes the maximum (minimum) over the elements sharing the vertex; then
// Hanging_vertex_matcher::match (Hanging_vertex_matcher.cpp:13-41) interpolates the hanging vertices of every hanging
// face from its corners.
void share_vertex(const hx_host_mesh* m, const Vertex_data& V, std::vector<double>& vc, bool minimum)
{
  const int nd = m->n_dim, nvert = 1 << nd;
  const double inf = std::numeric_limits<double>::infinity();
  std::vector<double> shared(V.n_vertex, minimum ? inf : -inf);
  for (size_t k = 0; k < V.vid.size(); ++k)
    shared[V.vid[k]] = minimum ? std::min(shared[V.vid[k]], vc[k]) : std::max(shared[V.vid[k]], vc[k]);
  for (size_t k = 0; k < V.vid.size(); ++k) vc[k] = shared[V.vid[k]];
  const std::vector<double> interp_mat{1., 0., .5, .5, 0., 1.};  // 3 x 2
  const size_t n_match = V.hanging.size() / 12;
  for (size_t i = 0; i < n_match; ++i) {
    const int* h = &V.hanging[12 * i];
    const int id = h[0], isp = h[1];
    const int str[2] = {h[2], h[3]};
    int n_fine = 1 << (nd - 1);
    for (int d = 0; d < nd - 1; ++d)
      if (str[d]) n_fine /= 2;
    int elem[4];
    for (int f = 0; f < n_fine; ++f) elem[f] = (h[4 + 2 * f] ? m->n_elem[0] : 0) + h[5 + 2 * f];
    const int n_vert = 1 << (nd - 1), stride = 1 << (nd - 1 - id);
    int inds[4];
    std::vector<double> values(n_vert);
    for (int iv = 0; iv < n_vert; ++iv) {
      inds[iv] = iv / stride * stride * 2 + iv % stride + isp * stride;
      values[iv] = vc[size_t(elem[stretched_ind(nd, iv, str)]) * nvert + inds[iv]];
    }
    const std::vector<double> interpolated = hypercube_matvec(interp_mat, 3, 2, values);
    for (int ie = 0; ie < n_fine; ++ie)
      for (int iv = 0; iv < n_vert; ++iv) {
        int interp_ind = (ie * !str[nd - 2]) % 2 + (iv % 2) * (1 + str[nd - 2]);
        if (nd == 3) interp_ind += ((ie * !str[0]) / (1 + !str[1]) + iv / 2 * (1 + str[0])) * 3;
        vc[size_t(elem[ie]) * nvert + inds[iv]] = interpolated[interp_ind];
      }
  }
}

// The fix coefficient of Solver::fix_admissibility (Solver.cpp:999-1030): element flags on the vertices, shared by
// maximum, the element maximum on all its vertices, shared again, interpolated n-linearly into the Laplacian
// artificial viscosity; returns the vertex values [n_elem[0] + n_elem[1]][nvert]
std::vector<double> fix_coefficient(hx_host_mesh* m, const std::vector<int>& record)
{
  const int nd = m->n_dim, rs = m->row_size, nq = ipow(rs, nd), nvert = 1 << nd;
  const Vertex_data& V = g_vertices;
  if (!V.set || V.n_elem[0] != m->n_elem[0] || V.n_elem[1] != m->n_elem[1])
    throw std::runtime_error("oracle: fix_admissibility needs the mesh's vertices (oracle_set_vertices)");
  const size_t n_tot = size_t(m->n_elem[0]) + m->n_elem[1];
  std::vector<double> vc(n_tot * nvert);
  for (size_t E = 0; E < n_tot; ++E)
    for (int v = 0; v < nvert; ++v) vc[E * nvert + v] = record[E];
  share_vertex(m, V, vc, false);
  for (size_t E = 0; E < n_tot; ++E) {
    double max_fac = 0;
    for (int v = 0; v < nvert; ++v) max_fac = std::max(max_fac, vc[E * nvert + v]);
    for (int v = 0; v < nvert; ++v) vc[E * nvert + v] = max_fac;
  }
  share_vertex(m, V, vc, false);
  Basis basis(rs);
  std::vector<double> interp(size_t(rs) * 2);
  for (int i = 0; i < rs; ++i) {
    interp[2 * i] = 1. - basis.node[i];
    interp[2 * i + 1] = basis.node[i];
  }
  for (int c = 0; c < 2; ++c)
    for (int e = 0; e < m->n_elem[c]; ++e) {
      const size_t E = (c ? m->n_elem[0] : 0) + e;
      std::vector<double> vf(vc.begin() + E * nvert, vc.begin() + (E + 1) * nvert);
      const std::vector<double> lap = hypercube_matvec(interp, rs, 2, vf);
      std::copy(lap.begin(), lap.end(), m->av[c] + (size_t(e) * 2 + 1) * nq);
    }
  return vc;
}

// Stab_art_visc (stabilizing_art_visc.cpp:8-61): per element, the specific volume's share of energy in the top
// Legendre mode along each dimension, (sum over rows of (row . orth_{rs-1} w)^2 face_weight) / (n_dim ||1/rho||^2),
// mapped onto [0, 1] by a sine ramp of half width 0.5 centred at log10 = -4.25 log10(rs - 1), times
// (rs - 1) char_speed h. Returns [n_elem[0] + n_elem[1]] (Cartesian first).
std::vector<double> stab_art_visc(const hx_host_mesh* m, double char_speed)
{
  const int nd = m->n_dim, rs = m->row_size, nq = ipow(rs, nd), nfq = nq / rs;
  const Basis basis(rs);
  const double ramp_center = -4.25 * std::log(rs - 1) / std::log(10), half_width = 0.5;
  std::vector<double> proj(rs), qw(nq), fw(nfq);
  for (int i = 0; i < rs; ++i) proj[i] = basis.ortho[(rs - 1) * rs + i] * basis.weight[i];
  for (int q = 0; q < nq; ++q) {
    qw[q] = 1.;
    for (int d = 0; d < nd; ++d) qw[q] *= basis.weight[(q / ipow(rs, nd - 1 - d)) % rs];
  }
  for (int f = 0; f < nfq; ++f) {
    fw[f] = 1.;
    for (int d = 0; d < nd - 1; ++d) fw[f] *= basis.weight[(f / ipow(rs, nd - 2 - d)) % rs];
  }
  std::vector<double> out(size_t(m->n_elem[0]) + m->n_elem[1]);
  for (int c = 0; c < 2; ++c) {
#pragma omp parallel for
    for (int e = 0; e < m->n_elem[c]; ++e) {
      const double* state = m->state[c] + size_t(e) * (nd + 2) * nq;
      std::vector<double> iv(nq);
      double norm_sq = 0;
      for (int q = 0; q < nq; ++q) {
        iv[q] = 1. / state[nd * nq + q];
        norm_sq += iv[q] * iv[q] * qw[q];
      }
      double nonsmooth = 0;
      for (int i_dim = 0; i_dim < nd; ++i_dim) {
        const int stride = ipow(rs, nd - 1 - i_dim);
        for (int f = 0; f < nfq; ++f) {
          // Row_index (Row_index.hpp:14-64): row f along i_dim starts at (f / stride) * stride * rs + f % stride
          const int base = (f / stride) * stride * rs + f % stride;
          double dot = 0;
          for (int k = 0; k < rs; ++k) dot += iv[base + k * stride] * proj[k];
          nonsmooth += dot * dot * fw[f];
        }
      }
      nonsmooth /= norm_sq * nd;
      double indicator = std::log(nonsmooth) / std::log(10);
      if (indicator <= ramp_center - half_width) indicator = 0;
      else if (indicator < ramp_center + half_width)
        indicator = .5 * (1 + std::sin(M_PI * (indicator - ramp_center) / 2 / half_width));
      else indicator = 1;
      out[(c ? m->n_elem[0] : 0) + e] = (rs - 1) * char_speed * m->nom_size[c][e] * indicator;
    }
  }
  return out;
}

// Solver::set_art_visc_admis (Solver.cpp:636-659): Stab_art_visc's element values on every vertex, shared by maximum
// (hanging vertices matched), interpolated n-linearly into the Laplacian artificial viscosity; returns the vertex values
std::vector<double> set_art_visc_admis(hx_host_mesh* m, double char_speed, std::vector<double>* uncert)
{
  const int nd = m->n_dim, rs = m->row_size, nq = ipow(rs, nd), nvert = 1 << nd;
  const Vertex_data& V = g_vertices;
  if (!V.set || V.n_elem[0] != m->n_elem[0] || V.n_elem[1] != m->n_elem[1])
    throw std::runtime_error("oracle: set_art_visc_admis needs the mesh's vertices (oracle_set_vertices)");
  for (int c = 0; c < 2; ++c)
    if (m->n_elem[c] && !m->av[c]) throw std::runtime_error("oracle: set_art_visc_admis needs the av arrays");
  const std::vector<double> u = stab_art_visc(m, char_speed);
  if (uncert) *uncert = u;
  std::vector<double> vc(u.size() * nvert);
  for (size_t E = 0; E < u.size(); ++E)
    for (int v = 0; v < nvert; ++v) vc[E * nvert + v] = u[E];
  share_vertex(m, V, vc, false);
  // Gauss_lobatto(2).interpolate(nodes): the linear interpolant of the two end values (Gauss_lobatto.cpp)
  const Basis basis(rs);
  std::vector<double> interp(size_t(rs) * 2);
  for (int i = 0; i < rs; ++i) {
    interp[2 * i] = 1. - basis.node[i];
    interp[2 * i + 1] = basis.node[i];
  }
  for (int c = 0; c < 2; ++c)
    for (int e = 0; e < m->n_elem[c]; ++e) {
      const size_t E = (c ? m->n_elem[0] : 0) + e;
      const std::vector<double> vf(vc.begin() + E * nvert, vc.begin() + (E + 1) * nvert);
      const std::vector<double> lap = hypercube_matvec(interp, rs, 2, vf);
      std::copy(lap.begin(), lap.end(), m->av[c] + (size_t(e) * 2 + 1) * nq);
    }
  return vc;
}

// ---- mesh metrics (Solver::calc_jacobian, Solver.cpp:273-381) ------------------------------------------------------

// math::dimension_matvec (math.cpp:58-84): `mat` (rows x cols) applied along dimension i_dim (0 slowest) of a
// hypercubic array of side cols
std::vector<double> dimension_matvec(const double* mat, int rows, int cols, const std::vector<double>& vec, int i_dim)
{
  const int n_rows = ipow(cols, i_dim + 1);
  const int n_cols = int(vec.size()) / n_rows;
  std::vector<double> prod(vec.size() * rows / cols);
  for (int i_outer = 0; i_outer < n_rows / cols; ++i_outer)
    for (int r = 0; r < rows; ++r)
      for (int j = 0; j < n_cols; ++j) {
        double acc = 0.;
        for (int c = 0; c < cols; ++c) acc += mat[r * cols + c] * vec[(size_t(i_outer) * cols + c) * n_cols + j];
        prod[(size_t(i_outer) * rows + r) * n_cols + j] = acc;
      }
  return prod;
}

// determinant of an n x n row-major matrix, n <= 3 (cofactor expansion along the first row)
double determinant(const double* a, int n)
{
  if (n == 1) return a[0];
  if (n == 2) return a[0] * a[3] - a[1] * a[2];
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}

// Deformed_element::position (Deformed_element.cpp:15-58): n-linear interpolation of the vertex positions plus, along
// each reference direction, the face node adjustments interpolated between the two faces times the element's edge
// vector in that direction. vert [nvert][nd]; adj [2 nd][nfq] or null
void deformed_position(int nd, int rs, const Basis& basis, const double* vert, const double* adj, int q, double* pos)
{
  const int nvert = 1 << nd, nq = ipow(rs, nd), nfq = nq / rs;
  for (int i_dim = 0; i_dim < nd; ++i_dim) {
    std::vector<double> vp(nvert);
    for (int v = 0; v < nvert; ++v) vp[v] = vert[v * nd + i_dim];
    std::vector<std::vector<double>> adjustments(nd, vp);
    for (int j_dim = nd - 1; j_dim >= 0; --j_dim) {
      const int stride = ipow(rs, nd - j_dim - 1);
      const double node = basis.node[(q / stride) % rs];
      const double interp[2] = {1. - node, node};
      vp = dimension_matvec(interp, 1, 2, vp, j_dim);
      int i_face_qpoint = 0;
      for (int k_dim = 0, face_stride = nq / rs / rs; k_dim < nd; ++k_dim)
        if (k_dim != j_dim) {
          i_face_qpoint += ((q / ipow(rs, nd - k_dim - 1)) % rs) * face_stride;
          face_stride /= rs;
        }
      for (int k_dim = 0; k_dim < nd; ++k_dim) {
        double contract[2] = {interp[0], interp[1]};
        if (j_dim == k_dim) {
          // adjust.cwiseProduct(interp) * [[-1, 1], [-1, 1]]
          const int i_adjust = 2 * j_dim * nfq + i_face_qpoint;
          const double p0 = (adj ? adj[i_adjust] : 0.) * interp[0], p1 = (adj ? adj[i_adjust + nfq] : 0.) * interp[1];
          contract[0] = -p0 - p1;
          contract[1] = p0 + p1;
        }
        adjustments[k_dim] = dimension_matvec(contract, 1, 2, adjustments[k_dim], j_dim);
      }
    }
    pos[i_dim] = vp[0];
    for (int j_dim = 0; j_dim < nd; ++j_dim) pos[i_dim] += adjustments[j_dim][0];
  }
}

// Deformed_element::set_jacobian (Deformed_element.cpp:60-136) of deformed element e: metrics, the element's face
// normals (into its face state blocks) and its vertex time step scale (vtss [nvert])
void deformed_set_jacobian(hx_host_mesh* m, const Mesh_view& mv, const Basis& basis, int e, const double* vert,
                           const double* adj, double* vtss)
{
  const int nd = m->n_dim, rs = m->row_size, nq = ipow(rs, nd), nfq = nq / rs, nvert = 1 << nd;
  const double nom_sz = m->nom_size[HX_DEFORMED][e];
  std::vector<double> jac(size_t(nd) * nd * nq);
  std::vector<std::vector<double>> pos(nd, std::vector<double>(nq));
  for (int q = 0; q < nq; ++q) {
    double p[3];
    deformed_position(nd, rs, basis, vert, adj, q, p);
    for (int i = 0; i < nd; ++i) pos[i][q] = p[i];
  }
  for (int i = 0; i < nd; ++i)
    for (int j = 0; j < nd; ++j) {
      const std::vector<double> d = dimension_matvec(basis.diff, rs, rs, pos[i], j);
      for (int q = 0; q < nq; ++q) jac[(size_t(i) * nd + j) * nq + q] = d[q] / nom_sz;
    }
  double* jac_dat = m->metrics + size_t(e) * (nd * nd + 1) * nq;
  for (int q = 0; q < nq; ++q) {
    double qj[9];
    for (int i = 0; i < nd; ++i)
      for (int j = 0; j < nd; ++j) qj[i * nd + j] = jac[(size_t(i) * nd + j) * nq + q];
    jac_dat[size_t(nd) * nd * nq + q] = determinant(qj, nd);
    for (int i = 0; i < nd; ++i)
      for (int j = 0; j < nd; ++j) {
        double copy[9];
        std::copy(qj, qj + nd * nd, copy);
        for (int r = 0; r < nd; ++r) copy[r * nd + i] = r == j;
        jac_dat[(size_t(i) * nd + j) * nq + q] = determinant(copy, nd);
      }
  }
  for (int i_dim = 0; i_dim < nd; ++i_dim)
    for (int sign = 0; sign < 2; ++sign) {
      double* face_data = mv.slot(m->face_slot[HX_DEFORMED][e * 2 * nd + 2 * i_dim + sign]);
      std::vector<std::vector<double>> face_jac(nd * nd);
      for (int i_jac = 0; i_jac < nd * nd; ++i_jac)
        face_jac[i_jac] = dimension_matvec(basis.bound + sign * rs, 1, rs,
                                           std::vector<double>(jac.begin() + size_t(i_jac) * nq,
                                                               jac.begin() + size_t(i_jac + 1) * nq), i_dim);
      for (int fq = 0; fq < nfq; ++fq) {
        double qj[9];
        for (int j = 0; j < nd; ++j)
          for (int k = 0; k < nd; ++k) qj[j * nd + k] = face_jac[j * nd + k][fq];
        for (int j = 0; j < nd; ++j) {
          for (int r = 0; r < nd; ++r) qj[r * nd + i_dim] = r == j;
          face_data[j * nfq + fq] = determinant(qj, nd);
        }
      }
    }
  std::vector<double> bound2(basis.bound, basis.bound + 2 * rs);
  std::vector<std::vector<double>> vertex_nrml(nd * nd);
  for (int i_jac = 0; i_jac < nd * nd; ++i_jac)
    vertex_nrml[i_jac] = hypercube_matvec(bound2, 2, rs, std::vector<double>(jac_dat + size_t(i_jac) * nq,
                                                                              jac_dat + size_t(i_jac + 1) * nq));
  const std::vector<double> vertex_det =
      hypercube_matvec(bound2, 2, rs, std::vector<double>(jac_dat + size_t(nd) * nd * nq, jac_dat + size_t(nd * nd + 1) * nq));
  for (int v = 0; v < nvert; ++v) {
    double norm_sum = 0.;
    for (int i = 0; i < nd; ++i) {
      double norm_sq = 0.;
      for (int j = 0; j < nd; ++j) norm_sq += vertex_nrml[i * nd + j][v] * vertex_nrml[i * nd + j][v];
      norm_sum += std::sqrt(norm_sq);
    }
    vtss[v] = nom_sz * vertex_det[v] / norm_sum;
  }
}

// Solver::calc_jacobian (Solver.cpp:273-381) on hx_host_mesh, without snapping or deformed hanging faces
void calc_jacobian(hx_host_mesh* m, const hx_vertex_data* vd)
{
  const int nd = m->n_dim, rs = m->row_size, nq = ipow(rs, nd), nfq = nq / rs, nvert = 1 << nd;
  Basis basis(rs);
  Mesh_view mv{m};
  Vertex_data V;
  V.n_vertex = vd->n_vertex;
  V.n_elem[0] = m->n_elem[0];
  V.n_elem[1] = m->n_elem[1];
  for (int c = 0; c < 2; ++c)
    for (long long k = 0; k < (long long)m->n_elem[c] * nvert; ++k) {
      const int id = vd->vertex_id[c][k];
      if (id < 0 || id >= vd->n_vertex) throw std::runtime_error("oracle: vertex id out of range");
      V.vid.push_back(id);
    }
  V.hanging.assign(vd->hanging, vd->hanging + 12LL * vd->n_hanging);
  // element jacobians, Cartesian then deformed (Element.cpp:99-112, Deformed_element.cpp:60-136)
  std::vector<double> vtss(size_t(m->n_elem[0] + m->n_elem[1]) * nvert);
  for (int e = 0; e < m->n_elem[HX_CARTESIAN]; ++e) {
    for (int f = 0; f < 2 * nd; ++f) {
      double* face = mv.slot(m->face_slot[HX_CARTESIAN][e * 2 * nd + f]);
      for (int j = 0; j < nd; ++j)
        for (int q = 0; q < nfq; ++q) face[j * nfq + q] = f / 2 == j;
    }
    for (int v = 0; v < nvert; ++v) vtss[size_t(e) * nvert + v] = m->nom_size[HX_CARTESIAN][e] / nd;  // Element.cpp:17
  }
  for (int e = 0; e < m->n_elem[HX_DEFORMED]; ++e)
    deformed_set_jacobian(m, mv, basis, e, vd->position + size_t(e) * nvert * nd,
                          vd->node_adj ? vd->node_adj + size_t(e) * 2 * nd * nfq : nullptr,
                          vtss.data() + (size_t(m->n_elem[0]) + e) * nvert);
  // boundary ghost faces copy the inside normal
  for (int i = 0; i < m->n_bc_con; ++i) {
    const int con = m->bc_con[i];
    const double* in_f = mv.slot(m->con_slot[HX_DEFORMED][2 * con]);
    double* gh_f = mv.slot(m->con_slot[HX_DEFORMED][2 * con + 1]);
    for (int k = 0; k < nd * nfq; ++k) gh_f[k] = in_f[k];
  }
  // the shared face normal of every deformed connection (boundary connections included)
  for (int con = 0; con < m->n_con[HX_DEFORMED]; ++con) {
    Dir dir = unpack_dir(m->con_dir[HX_DEFORMED][con]);
    double* elem_nrml[2] = {mv.slot(m->con_slot[HX_DEFORMED][2 * con]), mv.slot(m->con_slot[HX_DEFORMED][2 * con + 1])};
    // face_permutation (kernels_convective.cpp:38-41) of the nd normal components of face 1
    dispatch<Run_face_permutation>(nd, rs, dir, elem_nrml[1], nd, true);
    int sign[2];
    for (int s_ = 0; s_ < 2; ++s_) sign[s_] = 1 - 2 * dir.flip_normal(s_);
    for (int k = 0; k < nd * nfq; ++k) {
      double n = 0.;
      for (int s_ = 0; s_ < 2; ++s_) n += .5 * sign[s_] * elem_nrml[s_][k];
      for (int s_ = 0; s_ < 2; ++s_) elem_nrml[s_][k] = sign[s_] * n;
    }
    dispatch<Run_face_permutation>(nd, rs, dir, elem_nrml[1], nd, false);
    const int ns0 = m->con_nslot[2 * con], ns1 = m->con_nslot[2 * con + 1];
    std::copy(elem_nrml[0], elem_nrml[0] + nd * nfq, mv.nslot(ns0));
    if (ns1 != ns0) std::copy(elem_nrml[1], elem_nrml[1] + nd * nfq, mv.nslot(ns1));
  }
  share_vertex(m, V, vtss, true);
  for (int c = 0; c < 2; ++c)
    for (int e = 0; e < m->n_elem[c]; ++e)
      for (int v = 0; v < nvert; ++v)
        m->vertex_tss[c][size_t(e) * nvert + v] = vtss[((c ? m->n_elem[0] : 0) + size_t(e)) * nvert + v];
}

void swap_av(hx_host_mesh* m)
{
  const int nq = ipow(m->row_size, m->n_dim);
  for (int c = 0; c < 2; ++c)
    for (int e = 0; e < m->n_elem[c]; ++e)
      for (int q = 0; q < nq; ++q) std::swap(m->av[c][(size_t(e) * 2 + 1) * nq + q], m->av[c][size_t(e) * 2 * nq + q]);
}

// the fix's boundary faces: ghost state = inside (Solver.cpp:1066-1071), ghost LDG = -inside (Solver.cpp:103-115)
void fta_bcs(hx_host_mesh* m, bool flux)
{
  Mesh_view mv{m};
  const int nfq = ipow(m->row_size, m->n_dim - 1), nv = m->n_dim + 2;
  for (int i = 0; i < m->n_bc_con; ++i) {
    const int con = m->bc_con[i];
    const double* in_f = mv.slot(m->con_slot[HX_DEFORMED][2 * con]);
    double* gh_f = mv.slot(m->con_slot[HX_DEFORMED][2 * con + 1]);
    for (int k = 0; k < nv * nfq; ++k) {
      if (flux) gh_f[nv * nfq + k] = -in_f[nv * nfq + k];
      else gh_f[k] = in_f[k];
    }
  }
}

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }
void oracle_set_threads(int n) { omp_set_num_threads(n); }
int oracle_get_threads(void) { return omp_get_max_threads(); }

int oracle_write_face(hx_host_mesh* m, int cls, const hx_pde* pde)
{
  return guarded([&] {
    check_mesh(m);
    n_extrap(m, pde);
    int kind = pde_kind(pde);
    if (kind == HX_PDE_ADVECTION || kind == HX_PDE_SMOOTH_AV || kind == HX_PDE_FIX_ADMIS)
      dispatch<Run_av_write_face>(m->n_dim, m->row_size, m, cls, kind);
    else dispatch<Run_write_face>(m->n_dim, m->row_size, m, cls);
  });
}

int oracle_neighbor(hx_host_mesh* m, int cls, const hx_pde* pde, int)
{
  return guarded([&] {
    check_mesh(m);
    int kind = pde_kind(pde);
    if (kind == HX_PDE_ADVECTION || kind == HX_PDE_SMOOTH_AV || kind == HX_PDE_FIX_ADMIS)
      dispatch<Run_av_neighbor>(m->n_dim, m->row_size, m, cls, kind);
    else dispatch<Run_neighbor>(m->n_dim, m->row_size, m, cls, pde);
  });
}

int oracle_local(hx_host_mesh* m, int cls, const hx_pde* pde, double dt, int i_stage, int compute_residual, int use_filter)
{
  return guarded([&] {
    check_mesh(m);
    // HEXED_ASSERTs of the Local constructor (Spatial.hpp:322-323)
    int kind = pde_kind(pde);
    if ((kind == HX_PDE_NAVIER_STOKES || kind == HX_PDE_FIX_ADMIS) && i_stage)
      throw std::runtime_error("two-stage stabilization is not applicable to diffusion equations");
    if (i_stage && compute_residual) throw std::runtime_error("residual calculation is a single-stage operation");
    if (kind == HX_PDE_FIX_ADMIS) {
      if (compute_residual || use_filter)
        throw std::runtime_error("oracle: Fix_therm_admis is restated without compute_residual and filter");
      dispatch<Run_fta_local>(m->n_dim, m->row_size, m, cls, dt);
      return;
    }
    if (kind == HX_PDE_ADVECTION || kind == HX_PDE_SMOOTH_AV) {
      if (use_filter) throw std::runtime_error("oracle: the artificial-viscosity PDEs are restated without the filter");
      if (kind == HX_PDE_ADVECTION)
        dispatch<Run_adv_local>(m->n_dim, m->row_size, m, cls, dt, i_stage, compute_residual, pde->advect_length);
      else {
        if (i_stage) throw std::runtime_error("two-stage stabilization is not applicable to diffusion equations");
        if (compute_residual) throw std::runtime_error("oracle: Smooth_art_visc is restated without compute_residual");
        dispatch<Run_sav_local>(m->n_dim, m->row_size, m, cls, dt, pde->diff_time);
      }
      return;
    }
    dispatch<Run_local>(m->n_dim, m->row_size, m, cls, pde, dt, i_stage, compute_residual, use_filter);
  });
}

int oracle_neighbor_reconcile(hx_host_mesh* m, int cls, const hx_pde* pde)
{
  return guarded([&] {
    check_mesh(m);
    if (pde_kind(pde) == HX_PDE_SMOOTH_AV) dispatch<Run_sav_neighbor_reconcile>(m->n_dim, m->row_size, m, cls);
    else dispatch<Run_neighbor_reconcile>(m->n_dim, m->row_size, m, cls);  // the nv LDG variables
  });
}

int oracle_reconcile_ldg_flux(hx_host_mesh* m, int cls, const hx_pde* pde, double dt, int i_stage, int compute_residual,
                              int use_filter)
{
  return guarded([&] {
    check_mesh(m);
    if (pde_kind(pde) == HX_PDE_SMOOTH_AV) {
      if (i_stage) throw std::runtime_error("two-stage stabilization is not applicable to diffusion equations");
      if (compute_residual || use_filter)
        throw std::runtime_error("oracle: Smooth_art_visc is restated without compute_residual and filter");
      dispatch<Run_sav_reconcile>(m->n_dim, m->row_size, m, cls, dt, pde->diff_time, pde->cheby_step);
      return;
    }
    if (pde_kind(pde) == HX_PDE_FIX_ADMIS) {
      if (i_stage) throw std::runtime_error("two-stage stabilization is not applicable to diffusion equations");
      if (compute_residual || use_filter)
        throw std::runtime_error("oracle: Fix_therm_admis is restated without compute_residual and filter");
      dispatch<Run_fta_reconcile>(m->n_dim, m->row_size, m, cls, dt);
      return;
    }
    if (is_ns(pde) && i_stage) throw std::runtime_error("two-stage stabilization is not applicable to diffusion equations");
    dispatch<Run_reconcile>(m->n_dim, m->row_size, m, cls, dt, compute_residual, use_filter);
  });
}

int oracle_max_dt(hx_host_mesh* m, int cls, const hx_pde* pde, int is_local, int, double safety_conv, double safety_diff,
                  double* dt_out)
{
  return guarded([&] {
    check_mesh(m);
    int kind = pde_kind(pde);
    if (kind == HX_PDE_ADVECTION || kind == HX_PDE_SMOOTH_AV || kind == HX_PDE_FIX_ADMIS)
      *dt_out = dispatch<Run_av_max_dt>(m->n_dim, m->row_size, m, cls, kind != HX_PDE_ADVECTION, is_local, safety_conv,
                                        safety_diff);
    else *dt_out = dispatch<Run_max_dt>(m->n_dim, m->row_size, m, cls, pde, is_local, safety_conv, safety_diff);
  });
}

int oracle_prolong_refined(hx_host_mesh* m, const hx_pde* pde, int scale, int offset)
{
  return guarded([&] {
    check_mesh(m);
    dispatch<Run_prolong>(m->n_dim, m->row_size, m, n_extrap(m, pde), scale, offset);
  });
}

int oracle_restrict_refined(hx_host_mesh* m, const hx_pde* pde, int scale, int offset)
{
  return guarded([&] {
    check_mesh(m);
    dispatch<Run_restrict>(m->n_dim, m->row_size, m, n_extrap(m, pde), scale, offset);
  });
}

int oracle_apply_state_bcs(hx_host_mesh* m) { return guarded([&] { apply_state_bcs(m); }); }
int oracle_apply_flux_bcs(hx_host_mesh* m) { return guarded([&] { apply_flux_bcs(m); }); }
int oracle_characteristics(int n_dim, const double* ref, const double* normal, const double* state, double* vals,
                           double* parts)
{
  return guarded([&] {
    if (n_dim < 1 || n_dim > 3) throw std::runtime_error("invalid dimensionality");
    const Char_decomp ch(n_dim, ref, normal);
    double p[5][3];
    ch.split(state, p);
    for (int c = 0; c < 3; ++c) vals[c] = ch.vals[c];
    for (int v = 0; v < n_dim + 2; ++v)
      for (int c = 0; c < 3; ++c) parts[v * 3 + c] = p[v][c];
  });
}

// compute_euler (kernels_convective.cpp:8-18)
int oracle_compute_euler(hx_host_mesh* m, const hx_kernel_options* o)
{
  return guarded([&] {
    check_mesh(m);
    hx_pde p = euler_pde();
    const int nd = m->n_dim, rs = m->row_size;
    dispatch<Run_neighbor>(nd, rs, m, HX_CARTESIAN, &p);
    dispatch<Run_neighbor>(nd, rs, m, HX_DEFORMED, &p);
    dispatch<Run_restrict>(nd, rs, m, nd + 2, 1, 0);
    if (o->i_stage && o->compute_residual) throw std::runtime_error("residual calculation is a single-stage operation");
    dispatch<Run_local>(nd, rs, m, HX_CARTESIAN, &p, o->dt, o->i_stage, o->compute_residual, o->use_filter);
    dispatch<Run_local>(nd, rs, m, HX_DEFORMED, &p, o->dt, o->i_stage, o->compute_residual, o->use_filter);
    dispatch<Run_prolong>(nd, rs, m, nd + 2, 0, 0);
  });
}

// compute_navier_stokes (kernels_diffusive.cpp:8-29), flux_bc = Solver::apply_flux_bcs
int oracle_compute_navier_stokes(hx_host_mesh* m, const hx_kernel_options* o, hx_transport visc, hx_transport cond)
{
  return guarded([&] {
    check_mesh(m);
    hx_pde p = ns_pde(visc, cond);
    const int nd = m->n_dim, rs = m->row_size;
    if (o->i_stage) throw std::runtime_error("two-stage stabilization is not applicable to diffusion equations");
    dispatch<Run_neighbor>(nd, rs, m, HX_CARTESIAN, &p);
    dispatch<Run_neighbor>(nd, rs, m, HX_DEFORMED, &p);
    dispatch<Run_restrict>(nd, rs, m, nd + 2, 1, 0);
    dispatch<Run_restrict>(nd, rs, m, nd + 2, 0, 1);
    dispatch<Run_local>(nd, rs, m, HX_CARTESIAN, &p, o->dt, o->i_stage, o->compute_residual, o->use_filter);
    dispatch<Run_local>(nd, rs, m, HX_DEFORMED, &p, o->dt, o->i_stage, o->compute_residual, o->use_filter);
    if (!o->i_stage) {
      dispatch<Run_prolong>(nd, rs, m, nd + 2, 1, 1);
      apply_flux_bcs(m);
      dispatch<Run_neighbor_reconcile>(nd, rs, m, HX_CARTESIAN);
      dispatch<Run_neighbor_reconcile>(nd, rs, m, HX_DEFORMED);
      dispatch<Run_restrict>(nd, rs, m, nd + 2, 1, 1);
      dispatch<Run_reconcile>(nd, rs, m, HX_CARTESIAN, o->dt, o->compute_residual, o->use_filter);
      dispatch<Run_reconcile>(nd, rs, m, HX_DEFORMED, o->dt, o->compute_residual, o->use_filter);
    }
    dispatch<Run_prolong>(nd, rs, m, nd + 2, 0, 0);
  });
}

// max_dt_euler / max_dt_navier_stokes (kernels_max_dt.cpp:8-17)
int oracle_max_dt_euler(hx_host_mesh* m, const hx_kernel_options*, double msc, double msd, int local_time, double* dt_out)
{
  return guarded([&] {
    check_mesh(m);
    hx_pde p = euler_pde();
    double a = dispatch<Run_max_dt>(m->n_dim, m->row_size, m, HX_CARTESIAN, &p, local_time, msc, msd);
    double b = dispatch<Run_max_dt>(m->n_dim, m->row_size, m, HX_DEFORMED, &p, local_time, msc, msd);
    *dt_out = std::min(a, b);
  });
}

int oracle_max_dt_navier_stokes(hx_host_mesh* m, const hx_kernel_options*, double msc, double msd, int local_time,
                                hx_transport visc, hx_transport cond, double* dt_out)
{
  return guarded([&] {
    check_mesh(m);
    hx_pde p = ns_pde(visc, cond);
    double a = dispatch<Run_max_dt>(m->n_dim, m->row_size, m, HX_CARTESIAN, &p, local_time, msc, msd);
    double b = dispatch<Run_max_dt>(m->n_dim, m->row_size, m, HX_DEFORMED, &p, local_time, msc, msd);
    *dt_out = std::min(a, b);
  });
}

// Solver::update (Solver.cpp:834-886) without fix_admissibility (off by default, Solver.cpp:143)
int oracle_update(hx_host_mesh* m, const hx_update_options* o, hx_update_status* st)
{
  for (int i_flow = 0; i_flow < o->flow_iters; ++i_flow) {
    double safety = o->max_safety;
    double n_cheby = o->n_cheby_flow;
    double max_cheby = chebyshev_step(n_cheby, n_cheby - 1);
    for (int i_cheby = 0; i_cheby < n_cheby; ++i_cheby) {
      hx_kernel_options ko{0., 0, 0, o->use_filter};
      double mdt;
      int rc = o->use_ldg ? oracle_max_dt_navier_stokes(m, &ko, safety / max_cheby, safety, o->local_time, o->visc, o->cond, &mdt)
                          : oracle_max_dt_euler(m, &ko, safety / max_cheby, safety, o->local_time, &mdt);
      if (rc) return rc;
      double nominal_dt = std::min(mdt, o->max_time_step);
      double dt = nominal_dt * chebyshev_step(n_cheby, i_cheby);
      bool fixed = false;
      for (int i = 0; i < 2; ++i) {
        hx_kernel_options opts{dt, i, 0, o->use_filter};
        if ((rc = oracle_apply_state_bcs(m))) return rc;
        if (o->use_ldg && !i) rc = oracle_compute_navier_stokes(m, &opts, o->visc, o->cond);
        else rc = oracle_compute_euler(m, &opts);
        if (rc) return rc;
        if (o->fix_admis) {  // Solver.cpp:867
          int iters = 0;
          if ((rc = oracle_fix_admissibility(m, o, o->fix_admis_max_safety, &iters))) return rc;
          st->fix_admis_iters += iters;
          fixed = iters || fixed;
        }
      }
      if (fixed) break;
      st->time_step = dt;
      st->flow_time += dt;
    }
  }
  ++st->iteration;
  return 0;
}

// ---- thermodynamic admissibility ------------------------------------------------------------------------------

int oracle_set_vertices(const hx_host_mesh* m, int n_vertex, const int* vid_cartesian, const int* vid_deformed,
                        int n_hanging, const int* hanging)
{
  return guarded([&] {
    const int nd = m->n_dim, nvert = 1 << nd;
    Vertex_data v;
    v.n_elem[0] = m->n_elem[0];
    v.n_elem[1] = m->n_elem[1];
    v.n_vertex = n_vertex;
    for (int c = 0; c < 2; ++c) {
      const int* src = c ? vid_deformed : vid_cartesian;
      for (long long k = 0; k < (long long)m->n_elem[c] * nvert; ++k) {
        if (src[k] < 0 || src[k] >= n_vertex) throw std::runtime_error("oracle: vertex id out of range");
        v.vid.push_back(src[k]);
      }
    }
    if (n_hanging && nd < 2) throw std::runtime_error("oracle: hanging vertices need n_dim >= 2");
    for (long long k = 0; k < 12LL * n_hanging; ++k) v.hanging.push_back(hanging[k]);
    for (int i = 0; i < n_hanging; ++i) {
      const int* h = hanging + 12 * i;
      int n_fine = 1 << (nd - 1);
      for (int d = 0; d < nd - 1; ++d)
        if (h[2 + d]) n_fine /= 2;
      if (h[0] < 0 || h[0] >= nd) throw std::runtime_error("oracle: hanging-vertex dimension out of range");
      for (int f = 0; f < n_fine; ++f)
        if (h[4 + 2 * f] < 0 || h[4 + 2 * f] > 1 || h[5 + 2 * f] < 0 || h[5 + 2 * f] >= m->n_elem[h[4 + 2 * f]])
          throw std::runtime_error("oracle: hanging-vertex element out of range");
    }
    v.set = true;
    g_vertices = std::move(v);
  });
}

int oracle_calc_jacobian(hx_host_mesh* m, const hx_vertex_data* vertices)
{
  return guarded([&] {
    check_mesh(m);
    if (m->n_elem[HX_DEFORMED] && !vertices->position) throw std::runtime_error("oracle: vertex positions missing");
    calc_jacobian(m, vertices);
  });
}

int oracle_is_admissible(hx_host_mesh* m, int* admissible_out, int* record)
{
  return guarded([&] {
    check_mesh(m);
    std::vector<int> rec;
    *admissible_out = is_admissible(m, rec);
    if (record) std::copy(rec.begin(), rec.end(), record);
  });
}

// the fix coefficient alone (Solver.cpp:999-1030), from the given element flags, into the Laplacian coefficient;
// vertex_out (may be NULL) receives the vertex values [n_elem[0] + n_elem[1]][2^n_dim]
int oracle_fix_coefficient(hx_host_mesh* m, const int* record, double* vertex_out)
{
  return guarded([&] {
    check_mesh(m);
    for (int c = 0; c < 2; ++c)
      if (m->n_elem[c] && !m->av[c]) throw std::runtime_error("oracle: the fix coefficient needs the av arrays");
    std::vector<int> rec(record, record + m->n_elem[0] + m->n_elem[1]);
    const std::vector<double> vc = fix_coefficient(m, rec);
    if (vertex_out) std::copy(vc.begin(), vc.end(), vertex_out);
  });
}

int oracle_stab_art_visc(hx_host_mesh* m, double char_speed, double* uncert)
{
  return guarded([&] {
    check_mesh(m);
    const std::vector<double> u = stab_art_visc(m, char_speed);
    std::copy(u.begin(), u.end(), uncert);
  });
}

int oracle_set_art_visc_admis(hx_host_mesh* m, double char_speed, double* uncert, double* vertex_out)
{
  return guarded([&] {
    check_mesh(m);
    std::vector<double> u;
    const std::vector<double> vc = set_art_visc_admis(m, char_speed, &u);
    if (uncert) std::copy(u.begin(), u.end(), uncert);
    if (vertex_out) std::copy(vc.begin(), vc.end(), vertex_out);
  });
}

// compute_fix_therm_admis (kernels_diffusive.cpp:32), flux_bc = Solver::apply_fta_flux_bcs (Solver.cpp:103-115)
int oracle_compute_fix_therm_admis(hx_host_mesh* m, const hx_kernel_options* o)
{
  return guarded([&] {
    check_mesh(m);
    const int nd = m->n_dim, rs = m->row_size, nv = nd + 2;
    if (o->i_stage) throw std::runtime_error("two-stage stabilization is not applicable to diffusion equations");
    dispatch<Run_av_neighbor>(nd, rs, m, HX_CARTESIAN, int(HX_PDE_FIX_ADMIS));
    dispatch<Run_av_neighbor>(nd, rs, m, HX_DEFORMED, int(HX_PDE_FIX_ADMIS));
    dispatch<Run_restrict>(nd, rs, m, nv, 1, 0);
    dispatch<Run_restrict>(nd, rs, m, nv, 0, 1);
    dispatch<Run_fta_local>(nd, rs, m, HX_CARTESIAN, o->dt);
    dispatch<Run_fta_local>(nd, rs, m, HX_DEFORMED, o->dt);
    dispatch<Run_prolong>(nd, rs, m, nv, 1, 1);
    fta_bcs(m, true);
    dispatch<Run_neighbor_reconcile>(nd, rs, m, HX_CARTESIAN);
    dispatch<Run_neighbor_reconcile>(nd, rs, m, HX_DEFORMED);
    dispatch<Run_restrict>(nd, rs, m, nv, 1, 1);
    dispatch<Run_fta_reconcile>(nd, rs, m, HX_CARTESIAN, o->dt);
    dispatch<Run_fta_reconcile>(nd, rs, m, HX_DEFORMED, o->dt);
    dispatch<Run_prolong>(nd, rs, m, nv, 0, 0);
  });
}

// max_dt_fix_therm_admis (kernels_max_dt.cpp:21)
int oracle_max_dt_fix_therm_admis(hx_host_mesh* m, const hx_kernel_options*, double msc, double msd, int local_time,
                                  double* dt_out)
{
  return guarded([&] {
    check_mesh(m);
    const int nd = m->n_dim, rs = m->row_size;
    const double a = dispatch<Run_av_max_dt>(nd, rs, m, HX_CARTESIAN, true, local_time, msc, msd);
    const double b = dispatch<Run_av_max_dt>(nd, rs, m, HX_DEFORMED, true, local_time, msc, msd);
    *dt_out = std::min(a, b);
  });
}

// Solver::fix_admissibility (Solver.cpp:961-1095), step for step; the flow's Max_dt uses the settings `o`
int fix_admissibility(hx_host_mesh* m, const hx_update_options* o, double stability_ratio)
{
  for (int c = 0; c < 2; ++c)
    if (m->n_elem[c] && !m->av[c]) throw std::runtime_error("oracle: fix_admissibility needs the av arrays");
  std::vector<int> record;
  int iter;
  int n_iters = std::numeric_limits<int>::max();
  for (iter = 0; iter < n_iters; ++iter) {
    if (iter >= 100000)
      throw std::runtime_error("failed to fix thermodynamic admissability in " + std::to_string(iter) + " iterations");
    if (is_admissible(m, record)) {
      if (iter) n_iters = std::min(n_iters, 2 * iter);
      else {
        ++iter;
        break;
      }
    } else {
      n_iters = std::numeric_limits<int>::max();
    }
    fix_coefficient(m, record);
    swap_av(m);
    hx_kernel_options ko{0., 0, 0, 0};
    double dt = stability_ratio;
    if (oracle_max_dt_fix_therm_admis(m, &ko, dt, dt, 1, &dt)) throw std::runtime_error(g_err);
    dt = 1.;
    const double linear = dt, quadratic = dt * dt / 8 / 0.9;
    double step[2];
    step[1] = (linear + std::sqrt(linear * linear - 4 * quadratic)) / 2.;
    step[0] = quadratic / step[1];
    for (double st : step) {
      fta_bcs(m, false);
      ko.dt = st;
      if (oracle_compute_fix_therm_admis(m, &ko)) throw std::runtime_error(g_err);
    }
    const double n_cheby = o->n_cheby_flow, max_cheby = chebyshev_step(n_cheby, n_cheby - 1);
    hx_kernel_options fo{0., 0, 0, o->use_filter};
    double mdt;
    const int rc = o->use_ldg ? oracle_max_dt_navier_stokes(m, &fo, o->max_safety / max_cheby, o->max_safety, o->local_time,
                                                            o->visc, o->cond, &mdt)
                              : oracle_max_dt_euler(m, &fo, o->max_safety / max_cheby, o->max_safety, o->local_time, &mdt);
    if (rc) throw std::runtime_error(g_err);
    swap_av(m);
  }
  --iter;
  return iter;
}

int oracle_fix_admissibility(hx_host_mesh* m, const hx_update_options* o, double stability_ratio, int* n_iters)
{
  return guarded([&] {
    check_mesh(m);
    *n_iters = fix_admissibility(m, o, stability_ratio);
  });
}

// ---- artificial-viscosity generator: kernel sequences and Solver::update_art_visc_smoothness ------------------

int oracle_apply_advection_bcs(hx_host_mesh* m) { return guarded([&] { apply_advection_bcs(m); }); }
int oracle_apply_avc_diff_bcs(hx_host_mesh* m) { return guarded([&] { apply_avc_diff_bcs(m, false); }); }
int oracle_apply_avc_diff_flux_bcs(hx_host_mesh* m) { return guarded([&] { apply_avc_diff_bcs(m, true); }); }

// compute_advection (kernels_convective.cpp:8-19 with pde::Advection)
int oracle_compute_advection(hx_host_mesh* m, const hx_kernel_options* o, double advect_length)
{
  return guarded([&] {
    check_mesh(m);
    const int nd = m->n_dim, rs = m->row_size;
    dispatch<Run_av_neighbor>(nd, rs, m, HX_CARTESIAN, int(HX_PDE_ADVECTION));
    dispatch<Run_av_neighbor>(nd, rs, m, HX_DEFORMED, int(HX_PDE_ADVECTION));
    dispatch<Run_restrict>(nd, rs, m, nd + rs, 1, 0);
    dispatch<Run_adv_local>(nd, rs, m, HX_CARTESIAN, o->dt, o->i_stage, o->compute_residual, advect_length);
    dispatch<Run_adv_local>(nd, rs, m, HX_DEFORMED, o->dt, o->i_stage, o->compute_residual, advect_length);
    dispatch<Run_prolong>(nd, rs, m, nd + rs, 0, 0);
  });
}

// compute_smooth_av (kernels_diffusive.cpp:8-31 with pde::Smooth_art_visc), flux_bc = apply_avc_diff_flux_bcs
int oracle_compute_smooth_av(hx_host_mesh* m, const hx_kernel_options* o, double diff_time, double cheby_step)
{
  return guarded([&] {
    check_mesh(m);
    const int nd = m->n_dim, rs = m->row_size;
    if (o->i_stage) throw std::runtime_error("two-stage stabilization is not applicable to diffusion equations");
    dispatch<Run_av_neighbor>(nd, rs, m, HX_CARTESIAN, int(HX_PDE_SMOOTH_AV));
    dispatch<Run_av_neighbor>(nd, rs, m, HX_DEFORMED, int(HX_PDE_SMOOTH_AV));
    dispatch<Run_restrict>(nd, rs, m, 3, 1, 0);
    dispatch<Run_restrict>(nd, rs, m, 3, 0, 1);
    dispatch<Run_sav_local>(nd, rs, m, HX_CARTESIAN, o->dt, diff_time);
    dispatch<Run_sav_local>(nd, rs, m, HX_DEFORMED, o->dt, diff_time);
    dispatch<Run_prolong>(nd, rs, m, 3, 1, 1);
    apply_avc_diff_bcs(m, true);
    dispatch<Run_sav_neighbor_reconcile>(nd, rs, m, HX_CARTESIAN);
    dispatch<Run_sav_neighbor_reconcile>(nd, rs, m, HX_DEFORMED);
    dispatch<Run_restrict>(nd, rs, m, 3, 1, 1);
    dispatch<Run_sav_reconcile>(nd, rs, m, HX_CARTESIAN, o->dt, diff_time, cheby_step);
    dispatch<Run_sav_reconcile>(nd, rs, m, HX_DEFORMED, o->dt, diff_time, cheby_step);
    dispatch<Run_prolong>(nd, rs, m, 3, 0, 0);
  });
}

// Solver::update_art_visc_smoothness (Solver.cpp:457-582) with diffuse_art_visc (Solver.cpp:428-455)
int oracle_update_art_visc_smoothness(hx_host_mesh* m, const hx_av_options* o, double* residual)
{
  return guarded([&] {
    check_mesh(m);
    for (int c = 0; c < 2; ++c)
      if (m->n_elem[c] && (!m->forcing[c] || !m->advection[c] || !m->av[c]))
        throw std::runtime_error("artificial viscosity needs the av, forcing and advection arrays");
    const int nd = m->n_dim, rs = m->row_size;
    const double L = o->advect_length;
    dispatch<Run_av_points>(nd, rs, m, 0, o);
    // enforce CFL condition
    for (int c = 0; c < 2; ++c) dispatch<Run_write_face>(nd, rs, m, c);
    dispatch<Run_prolong>(nd, rs, m, nd + 2, 0, 0);
    dispatch<Run_av_max_dt>(nd, rs, m, HX_CARTESIAN, false, 1, o->advect_max_safety, 1.);
    dispatch<Run_av_max_dt>(nd, rs, m, HX_DEFORMED, false, 1, o->advect_max_safety, 1.);
    for (int iter = 0; iter < o->advect_iters; ++iter) {
      for (int c = 0; c < 2; ++c) dispatch<Run_av_write_face>(nd, rs, m, c, int(HX_PDE_ADVECTION));
      dispatch<Run_prolong>(nd, rs, m, nd + rs, 0, 0);
      for (int i = 0; i < 2; ++i) {
        apply_advection_bcs(m);
        hx_kernel_options ko{1., i, 0, 0};
        if (oracle_compute_advection(m, &ko, L)) throw std::runtime_error(g_err);
      }
    }
    dispatch<Run_av_points>(nd, rs, m, 1, o);
    // root-smear-square: diffuse_art_visc
    const int n_real = 3;
    const double diff_time = o->diff_ratio * L * L / n_real;
    dispatch<Run_av_max_dt>(nd, rs, m, HX_CARTESIAN, true, 1, 1., o->diff_max_safety);
    dispatch<Run_av_max_dt>(nd, rs, m, HX_DEFORMED, true, 1, 1., o->diff_max_safety);
    for (int c = 0; c < 2; ++c) dispatch<Run_av_write_face>(nd, rs, m, c, int(HX_PDE_SMOOTH_AV));
    dispatch<Run_prolong>(nd, rs, m, nd + 2, 0, 0);
    for (int iter = 0; iter < o->diff_iters; ++iter)
      for (int i_cheby = 0; i_cheby < o->n_cheby; ++i_cheby) {
        const double s = chebyshev_step(o->n_cheby, i_cheby);
        apply_avc_diff_bcs(m, false);
        hx_kernel_options ko{s, 0, 0, 0};
        if (oracle_compute_smooth_av(m, &ko, diff_time, s)) throw std::runtime_error(g_err);
      }
    const double resid = dispatch<Run_av_points>(nd, rs, m, 2, o);
    if (residual) *residual = std::sqrt(resid);
    for (int c = 0; c < 2; ++c) dispatch<Run_write_face>(nd, rs, m, c);
    dispatch<Run_prolong>(nd, rs, m, nd + 2, 0, 0);
  });
}

// Solver::compute_residual (Solver.cpp:899-913)
int oracle_compute_residual(hx_host_mesh* m, const hx_update_options* o)
{
  int rc = oracle_apply_state_bcs(m);
  if (rc) return rc;
  hx_kernel_options opts{1., 0, 1, o->use_filter};
  if (o->use_ldg) return oracle_compute_navier_stokes(m, &opts, o->visc, o->cond);
  return oracle_compute_euler(m, &opts);
}

}  // extern "C"
