// gfx950 kernels of Solver::calc_jacobian (Solver.cpp:273-381): the deformed elements' metrics, face normals and vertex
// time step scale (Deformed_element::set_jacobian and ::position, Deformed_element.cpp:15-136), the Cartesian
// elements' unit face normals (Element::set_jacobian, Element.cpp:99-112), the shared normal of every deformed
// connection, and the vertex time step scale shared by minimum with hanging vertices matched (Solver.cpp:40-54,
// Hanging_vertex_matcher.cpp:13-41). Included by hx_kernels.hpp.
#pragma once

namespace hx {

// basis constants, passed by value
template <int RS>
struct Metric_consts
{
  double node[RS];
  double diff[RS][RS];   // diff[i][j]: derivative of Lagrange polynomial j at node i
  double bound[2][RS];   // bound[s][j]: value of Lagrange polynomial j at face s
};

__device__ inline double det_n(const double* a, int n)
{
  if (n == 1) return a[0];
  if (n == 2) return a[0] * a[3] - a[1] * a[2];
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}

// determinant of the nd x nd matrix `a` with column `col` replaced by unit vector `row`
template <int ND>
__device__ inline double det_unit_col(const double* a, int col, int row)
{
  double c[ND * ND];
  for (int r = 0; r < ND; ++r)
    for (int k = 0; k < ND; ++k) c[r * ND + k] = (k == col) ? double(r == row) : a[r * ND + k];
  return det_n(c, ND);
}

template <int ND, int RS>
struct Jac_lds
{
  using S = Shape<ND, RS>;
  static constexpr int NQP = S::NQP, NVERT = 1 << ND, NM = ND * ND + 1;
  static constexpr int POS = 0;                   // [ND][NQP] quadrature point positions
  static constexpr int JAC = POS + ND * NQP;      // [ND*ND][NQP] Jacobian
  static constexpr int MET = JAC + ND * ND * NQP; // [ND*ND + 1][NQP] reference-level normals and determinant
  static constexpr int VX = MET + NM * NQP;       // [NM][NVERT] metrics extrapolated to the vertices
  static constexpr int PER = VX + NM * NVERT;
};

// Deformed_element::set_jacobian of every deformed element. vert: [n][nvert][nd]; adj: [n][2 nd][nfq] or null.
// Writes the metrics, the element's face normals into its face state blocks and its vertex time step scale into
// vtss[(n_cartesian + e) * nvert + v].
template <int ND, int RS>
__global__ __launch_bounds__((Shape<ND, RS>::BLOCK)) void k_set_jacobian(Dev_mesh m, Metric_consts<RS> mc,
                                                                         const double* vert, const double* adj,
                                                                         double* vtss)
{
  using S = Shape<ND, RS>;
  using R = Row<ND, RS>;
  using L = Jac_lds<ND, RS>;
  constexpr int NQ = S::NQ, NQP = S::NQP, NFQ = S::NFQ, EPB = S::EPB, NVERT = L::NVERT, NM = L::NM;
  __shared__ double lds[EPB * L::PER];
  const int t = threadIdx.x, el = t / NQ, q = t % NQ, qp = S::pad(q);
  const int n = m.n_elem[HX_DEFORMED];
  const int e = blockIdx.x * EPB + el;
  const bool valid = e < n;
  double* my = lds + el * L::PER;
  const double* vp = vert + size_t(valid ? e : 0) * NVERT * ND;
  const double* ap = (adj && valid) ? adj + size_t(e) * 2 * ND * NFQ : nullptr;
  // position (Deformed_element.cpp:15-58): the n-linear interpolation of the vertices plus, along each direction k,
  // the face adjustments interpolated between the faces (k, 0) and (k, 1) times the edge vector along k
  double x[ND];
  for (int d = 0; d < ND; ++d) x[d] = mc.node[R::node(q, d)];
  for (int i = 0; i < ND; ++i) {
    double p = 0.;
    for (int v = 0; v < NVERT; ++v) {
      double w = 1.;
      for (int d = 0; d < ND; ++d) w *= ((v >> (ND - 1 - d)) & 1) ? x[d] : 1. - x[d];
      p += w * vp[v * ND + i];
    }
    if (ap) {
      for (int k = 0; k < ND; ++k) {
        const int f = R::fq(q, k);
        const double a = ap[2 * k * NFQ + f] * (1. - x[k]) + ap[(2 * k + 1) * NFQ + f] * x[k];
        double edge = 0.;
        for (int v = 0; v < NVERT; ++v) {
          double w = 1.;
          for (int d = 0; d < ND; ++d)
            if (d != k) w *= ((v >> (ND - 1 - d)) & 1) ? x[d] : 1. - x[d];
          edge += (((v >> (ND - 1 - k)) & 1) ? w : -w) * vp[v * ND + i];
        }
        p += a * edge;
      }
    }
    my[L::POS + i * NQP + qp] = valid ? p : 0.;
  }
  __syncthreads();
  // Jacobian: derivative along each reference direction over the nominal size, then determinant and cofactors
  const double inv_h = valid ? 1. / m.nom_size[HX_DEFORMED][e] : 1.;
  double J[ND * ND];
  for (int j = 0; j < ND; ++j) {
    const int kq = R::node(q, j), base = R::base(q, j), st = R::stride(j);
    for (int i = 0; i < ND; ++i) {
      double acc = 0.;
      for (int k = 0; k < RS; ++k) acc += mc.diff[kq][k] * my[L::POS + i * NQP + S::pad(base + k * st)];
      J[i * ND + j] = acc * inv_h;
      my[L::JAC + (i * ND + j) * NQP + qp] = J[i * ND + j];
    }
  }
  double met[NM];
  met[ND * ND] = det_n(J, ND);
  for (int i = 0; i < ND; ++i)
    for (int j = 0; j < ND; ++j) met[i * ND + j] = det_unit_col<ND>(J, i, j);
  for (int k = 0; k < NM; ++k) my[L::MET + k * NQP + qp] = met[k];
  if (valid) {
    double* out = m.metrics + size_t(e) * NM * NQ;
    for (int k = 0; k < NM; ++k) out[k * NQ + q] = met[k];
  }
  __syncthreads();
  // face normals (Deformed_element.cpp:91-113): the Jacobian extrapolated to the face, column i_dim replaced by each
  // unit vector
  if (valid) {
    for (int k = q; k < 2 * ND * NFQ; k += NQ) {
      const int f = k / NFQ, fq = k % NFQ, i_dim = f / 2, sign = f % 2;
      const int base = R::base_of_fq(fq, i_dim), st = R::stride(i_dim);
      double FJ[ND * ND];
      for (int a = 0; a < ND * ND; ++a) {
        double acc = 0.;
        for (int r = 0; r < RS; ++r) acc += mc.bound[sign][r] * my[L::JAC + a * NQP + S::pad(base + r * st)];
        FJ[a] = acc;
      }
      double* face = slot_ptr(m, m.face_slot[HX_DEFORMED][e * 2 * ND + f]);
      for (int j = 0; j < ND; ++j) face[j * NFQ + fq] = det_unit_col<ND>(FJ, i_dim, j);
    }
  }
  // metrics extrapolated to the vertices (math::hypercube_matvec with the boundary matrix)
  for (int k = q; k < NM * NVERT; k += NQ) {
    const int a = k / NVERT, v = k % NVERT;
    double acc = 0.;
    for (int p = 0; p < NQ; ++p) {
      double w = 1.;
      for (int d = 0; d < ND; ++d) w *= mc.bound[(v >> (ND - 1 - d)) & 1][R::node(p, d)];
      acc += w * my[L::MET + a * NQP + S::pad(p)];
    }
    my[L::VX + a * NVERT + v] = acc;
  }
  __syncthreads();
  if (valid && q < NVERT) {
    double norm_sum = 0.;
    for (int i = 0; i < ND; ++i) {
      double sq = 0.;
      for (int j = 0; j < ND; ++j) sq += my[L::VX + (i * ND + j) * NVERT + q] * my[L::VX + (i * ND + j) * NVERT + q];
      norm_sum += sqrt(sq);
    }
    vtss[(size_t(m.n_elem[HX_CARTESIAN]) + e) * NVERT + q] =
        m.nom_size[HX_DEFORMED][e] * my[L::VX + ND * ND * NVERT + q] / norm_sum;
  }
}

// Element::set_jacobian of the Cartesian elements: unit normals on every face; vertex time step scale h / n_dim
// (Element.cpp:17)
template <int ND, int RS>
__global__ __launch_bounds__(256) void k_cartesian_jacobian(Dev_mesh m, double* vtss)
{
  constexpr int NFQ = Shape<ND, RS>::NFQ, NVERT = 1 << ND;
  const size_t gid = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (gid >= size_t(m.n_elem[HX_CARTESIAN]) * 2 * ND * NFQ) return;
  const int e = int(gid / (2 * ND * NFQ)), r = int(gid % (2 * ND * NFQ)), f = r / NFQ, fq = r % NFQ;
  double* face = slot_ptr(m, m.face_slot[HX_CARTESIAN][e * 2 * ND + f]);
  for (int j = 0; j < ND; ++j) face[j * NFQ + fq] = double(f / 2 == j);
  if (r == 0)
    for (int v = 0; v < NVERT; ++v) vtss[size_t(e) * NVERT + v] = m.nom_size[HX_CARTESIAN][e] / ND;
}

// boundary ghost faces copy the inside face normal (Solver.cpp:316-323)
template <int ND, int RS>
__global__ __launch_bounds__(256) void k_bc_copy_normal(Dev_mesh m)
{
  constexpr int NFQ = Shape<ND, RS>::NFQ;
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int i = gid / NFQ, f = gid % NFQ;
  if (i >= m.n_bc_con) return;
  const int con = m.bc_con[i];
  const double* in_f = slot_ptr(m, m.con_slot[HX_DEFORMED][2 * con]);
  double* gh_f = slot_ptr(m, m.con_slot[HX_DEFORMED][2 * con + 1]);
  for (int d = 0; d < ND; ++d) gh_f[d * NFQ + f] = in_f[d * NFQ + f];
}

// the shared normal of every deformed connection (Solver.cpp:324-352): face 1 permuted to face-0 order, the
// sign-corrected average of the two element normals written back to both faces and to the connection's normal slots
// (side 1 only when it has a slot of its own)
template <int ND, int RS>
__global__ __launch_bounds__(256) void k_connection_normals(Dev_mesh m)
{
  constexpr int NFQ = Shape<ND, RS>::NFQ;
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int con = gid / NFQ, f = gid % NFQ;
  if (con >= m.n_con[HX_DEFORMED]) return;
  const Dir dir(m.con_dir[HX_DEFORMED][con]);
  double* e0 = slot_ptr(m, m.con_slot[HX_DEFORMED][2 * con]);
  double* e1 = slot_ptr(m, m.con_slot[HX_DEFORMED][2 * con + 1]);
  const int p1 = face1_index<ND, RS>(dir, f);
  const double s0 = 1 - 2 * int(dir.flip_normal(0)), s1 = 1 - 2 * int(dir.flip_normal(1));
  const int ns0 = m.con_nslot[2 * con], ns1 = m.con_nslot[2 * con + 1];
  double* n0 = m.normals + size_t(ns0) * ND * NFQ;
  double* n1 = m.normals + size_t(ns1) * ND * NFQ;
  for (int d = 0; d < ND; ++d) {
    const double nv = .5 * s0 * e0[d * NFQ + f] + .5 * s1 * e1[d * NFQ + p1];
    e0[d * NFQ + f] = s0 * nv;
    e1[d * NFQ + p1] = s1 * nv;
    n0[d * NFQ + f] = s0 * nv;
    if (ns1 != ns0) n1[d * NFQ + p1] = s1 * nv;
  }
}

// Hanging_vertex_matcher::match (Hanging_vertex_matcher.cpp:13-41), one thread per hanging face, in two launches:
// mode 0 reads the corner values from the fine elements' copies and interpolates them into `vals` [n][9], mode 1
// writes them to the fine elements' copies. (No matcher reads a copy another one writes with a different value: the
// corners of a hanging face are vertices of its coarse level, the positions a matcher interpolates are not.)
template <int ND>
__global__ __launch_bounds__(256) void k_hanging_match(const int* hanging, int n_hanging, int n_cartesian,
                                                       double* vtss, double* vals, int mode)
{
  constexpr int NVERT = 1 << ND, NVF = 1 << (ND - 1);
  const int i = blockIdx.x * 256 + threadIdx.x;
  if constexpr (ND < 2) return;  // no hanging vertices in 1D (the host rejects them)
  else {
  if (i >= n_hanging) return;
  const int* h = hanging + 12 * i;
  const int id = h[0], isp = h[1];
  const int str[2] = {h[2], h[3]};
  int n_fine = NVF;
  for (int d = 0; d < ND - 1; ++d)
    if (str[d]) n_fine /= 2;
  int elem[4];
  for (int f = 0; f < n_fine; ++f) elem[f] = (h[4 + 2 * f] ? n_cartesian : 0) + h[5 + 2 * f];
  const int stride = 1 << (ND - 1 - id);
  int inds[NVF];
  for (int iv = 0; iv < NVF; ++iv) inds[iv] = iv / stride * stride * 2 + iv % stride + isp * stride;
  double* out = vals + size_t(i) * 9;
  if (mode == 0) {
    // math::stretched_ind (math.hpp:188-199) picks the fine element holding each corner
    double values[NVF];
    for (int iv = 0; iv < NVF; ++iv) {
      int st = 1, se = 0;
      for (int d = ND - 2; d >= 0; --d)
        if (!str[d]) {
          se += ((iv >> (ND - 2 - d)) & 1) * st;
          st *= 2;
        }
      values[iv] = vtss[size_t(elem[se]) * NVERT + inds[iv]];
    }
    // math::hypercube_matvec with {{1, 0}, {.5, .5}, {0, 1}}: 3^(nd-1) values at the corners and midpoints
    if constexpr (ND == 2) {
      out[0] = values[0];
      out[1] = .5 * values[0] + .5 * values[1];
      out[2] = values[1];
    } else if constexpr (ND == 3) {
      double fact[2][3];
      for (int r = 0; r < 2; ++r) {
        fact[r][0] = values[2 * r];
        fact[r][1] = .5 * values[2 * r] + .5 * values[2 * r + 1];
        fact[r][2] = values[2 * r + 1];
      }
      for (int c = 0; c < 3; ++c) {
        out[c] = fact[0][c];
        out[3 + c] = .5 * fact[0][c] + .5 * fact[1][c];
        out[6 + c] = fact[1][c];
      }
    }
    return;
  }
  for (int ie = 0; ie < n_fine; ++ie)
    for (int iv = 0; iv < NVF; ++iv) {
      int interp_ind = (ie * !str[ND - 2]) % 2 + (iv % 2) * (1 + str[ND - 2]);
      if constexpr (ND == 3) interp_ind += ((ie * !str[0]) / (1 + !str[1]) + iv / 2 * (1 + str[0])) * 3;
      vtss[size_t(elem[ie]) * NVERT + inds[iv]] = out[interp_ind];
    }
  }
}

}  // namespace hx
